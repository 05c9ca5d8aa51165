// host_io.cpp -- .egr binary CSR reader (memory-mapped, both nindex widths)
// and the crash-consistent wide writer.
//
// Format (README "File formats"; core/FlashGraph.h:108-169):
//   int32 nodes; int32 edges or -1; [int64 edges if -1];
//   nindex[nodes+1] int32 (narrow) or int64 (after the -1 sentinel);
//   nlist[edges] int32.
// The writer follows core/flash_compactor.h:40-104: write <path>.tmp.<pid>,
// fflush, fsync, rename over <path>, fsync the directory.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/misfit.h"

namespace {
struct EgrMap {
    void* base = nullptr;
    size_t size = 0;
    int fd = -1;
};

void put_err(char* err, int errlen, const std::string& m) {
    if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", m.c_str());
}
}  // namespace

extern "C" int misfit_egr_open(const char* path, misfit_egr* out, char* err, int errlen) {
    if (!path || !out) return MISFIT_E_ARG;
    memset(out, 0, sizeof(*out));
    int fd = ::open(path, O_RDONLY);
    if (fd < 0) {
        put_err(err, errlen, std::string("could not open ") + path);
        return MISFIT_E_IO;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        ::close(fd);
        put_err(err, errlen, std::string("fstat failed for ") + path);
        return MISFIT_E_IO;
    }
    size_t sz = (size_t)st.st_size;
    if (sz < 8) {
        ::close(fd);
        put_err(err, errlen, std::string("failed to read header from ") + path);
        return MISFIT_E_IO;
    }
    void* map = mmap(nullptr, sz, PROT_READ, MAP_PRIVATE, fd, 0);
    if (map == MAP_FAILED) {
        ::close(fd);
        put_err(err, errlen, std::string("mmap failed for ") + path);
        return MISFIT_E_IO;
    }
    const char* p = (const char*)map;
    int32_t nodes, e32;
    memcpy(&nodes, p, 4);
    memcpy(&e32, p + 4, 4);
    int64_t edges;
    size_t hdr;
    int wide;
    if (e32 == -1) {
        if (sz < 16) {
            munmap(map, sz);
            ::close(fd);
            put_err(err, errlen, std::string("failed to read 64-bit edge count from ") + path);
            return MISFIT_E_IO;
        }
        memcpy(&edges, p + 8, 8);
        hdr = 16;
        wide = 1;
    } else {
        edges = e32;
        hdr = 8;
        wide = 0;
    }
    if (nodes < 1 || edges < 0) {
        munmap(map, sz);
        ::close(fd);
        put_err(err, errlen, std::string("invalid header in ") + path);
        return MISFIT_E_IO;
    }
    size_t stride = wide ? 8 : 4;
    size_t need = hdr + ((size_t)nodes + 1) * stride + (size_t)edges * 4;
    if (sz < need) {
        munmap(map, sz);
        ::close(fd);
        put_err(err, errlen, std::string("truncated file ") + path);
        return MISFIT_E_IO;
    }
    // streamed once into device memory, so sequential read-ahead helps
    madvise(map, sz, MADV_SEQUENTIAL);
    EgrMap* m = new EgrMap{map, sz, fd};
    out->n = nodes;
    out->m = edges;
    out->wide = wide;
    out->nindex = p + hdr;
    out->nlist = (const int32_t*)(p + hdr + ((size_t)nodes + 1) * stride);
    out->priv = m;
    return MISFIT_OK;
}

extern "C" int misfit_egr_close(misfit_egr* g) {
    if (!g || !g->priv) return MISFIT_E_ARG;
    EgrMap* m = (EgrMap*)g->priv;
    munmap(m->base, m->size);
    ::close(m->fd);
    delete m;
    memset(g, 0, sizeof(*g));
    return MISFIT_OK;
}

extern "C" int misfit_write_egr(const char* path, int32_t n, int64_t m, const int64_t* nindex, const int32_t* nlist,
                                char* err, int errlen) {
    if (!path || n < 1 || m < 0 || !nindex || (m && !nlist)) return MISFIT_E_ARG;
    std::string out_path(path);
    std::string tmp = out_path + ".tmp." + std::to_string(getpid());
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) {
        put_err(err, errlen, "could not open " + tmp + " for writing");
        return MISFIT_E_IO;
    }
    const int32_t sentinel = -1;
    bool ok = fwrite(&n, 4, 1, f) == 1 && fwrite(&sentinel, 4, 1, f) == 1 && fwrite(&m, 8, 1, f) == 1 &&
              fwrite(nindex, 8, (size_t)n + 1, f) == (size_t)n + 1 &&
              (m == 0 || fwrite(nlist, 4, (size_t)m, f) == (size_t)m);
    if (!ok) {
        fclose(f);
        unlink(tmp.c_str());
        put_err(err, errlen, "write failed on " + tmp);
        return MISFIT_E_IO;
    }
    if (fflush(f) != 0 || fsync(fileno(f)) != 0) {
        fclose(f);
        unlink(tmp.c_str());
        put_err(err, errlen, "flush/fsync failed on " + tmp);
        return MISFIT_E_IO;
    }
    fclose(f);
    if (rename(tmp.c_str(), out_path.c_str()) != 0) {
        unlink(tmp.c_str());
        put_err(err, errlen, "rename(" + tmp + " -> " + out_path + ") failed");
        return MISFIT_E_IO;
    }
    std::string dir = ".";
    size_t slash = out_path.find_last_of('/');
    if (slash != std::string::npos) dir = slash == 0 ? "/" : out_path.substr(0, slash);
    int dfd = ::open(dir.c_str(), O_RDONLY);
    if (dfd >= 0) {
        fsync(dfd);
        ::close(dfd);
    }
    return MISFIT_OK;
}

// ---------------------------------------------------------------- .mtx -> .egr
// Two streaming passes over the text file, as tools/efficient_converter.py:
// 9-113 does: the first non-comment line is the dimension line (n = max of
// its first two fields); every later line "u v ..." with 1 <= u <= n puts
// v-1 into row u-1, in file order; lines with u > n are dropped; lines with
// fewer than two fields are skipped.  Nothing is symmetrised, sorted or
// deduplicated.  The output is the wide form, written through a mapping of
// the output file.  (The Python tool stops with an error on a '%' line after
// the header and on u < 1; here the first is skipped as a comment and the
// second is rejected with an error.)
namespace {
struct LineReader {
    FILE* f = nullptr;
    char* buf = nullptr;
    size_t cap = 0;
    ~LineReader() {
        if (f) fclose(f);
        free(buf);
    }
    bool open(const char* p) { return (f = fopen(p, "rb")) != nullptr; }
    ssize_t next() { return getline(&buf, &cap, f); }
};

// Two leading integer fields of a line; false when it has fewer than two.
bool two_fields(const char* s, long long& a, long long& b, int* nfields) {
    char* end = nullptr;
    int k = 0;
    long long v[3] = {0, 0, 0};
    const char* p = s;
    while (k < 3) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') p++;
        if (!*p) break;
        errno = 0;
        long long x = strtoll(p, &end, 10);
        if (end == p) {  // a non-numeric field: count it, stop reading numbers
            k++;
            break;
        }
        v[k++] = x;
        p = end;
        while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') p++;
    }
    if (nfields) *nfields = k;
    a = v[0];
    b = v[1];
    return k >= 2;
}
}  // namespace

extern "C" int misfit_mtx_to_egr(const char* mtx_path, const char* egr_path, int32_t* n_out, int64_t* m_out,
                                 char* err, int errlen) {
    if (!mtx_path || !egr_path) return MISFIT_E_ARG;
    LineReader r;
    if (!r.open(mtx_path)) {
        put_err(err, errlen, std::string("cannot open ") + mtx_path);
        return MISFIT_E_IO;
    }
    long long n = -1, a = 0, b = 0;
    ssize_t len;
    // header: first non-comment line
    while ((len = r.next()) >= 0) {
        if (r.buf[0] == '%') continue;
        int nf = 0;
        if (!two_fields(r.buf, a, b, &nf)) {
            put_err(err, errlen, "dimension line must hold two integers");
            return MISFIT_E_UNSUPPORTED;
        }
        n = a > b ? a : b;
        break;
    }
    if (n < 1 || n > INT32_MAX - 1) {
        put_err(err, errlen, "missing or invalid dimension line");
        return MISFIT_E_UNSUPPORTED;
    }
    const long data_pos = ftell(r.f);
    // pass 1: degrees
    std::string msg;
    int64_t* nindex = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (!nindex) return MISFIT_E_NOMEM;
    while ((len = r.next()) >= 0) {
        if (r.buf[0] == '%') continue;
        if (!two_fields(r.buf, a, b, nullptr)) continue;
        if (a < 1) {
            free(nindex);
            put_err(err, errlen, "vertex ids are 1-indexed: found " + std::to_string(a));
            return MISFIT_E_UNSUPPORTED;
        }
        if (a <= n) nindex[a]++;  // degree of row a-1 lands at nindex[a]
    }
    for (long long i = 1; i <= n; i++) nindex[i] += nindex[i - 1];
    const int64_t m = nindex[n];
    // output file: header, nindex, then nlist filled through a mapping
    const size_t hdr = 16 + ((size_t)n + 1) * 8;
    const size_t total = hdr + (size_t)m * 4;
    int fd = ::open(egr_path, O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) {
        free(nindex);
        put_err(err, errlen, std::string("cannot create ") + egr_path);
        return MISFIT_E_IO;
    }
    if (ftruncate(fd, (off_t)total) != 0) {
        ::close(fd);
        free(nindex);
        put_err(err, errlen, "cannot size the output file");
        return MISFIT_E_IO;
    }
    char* out = (char*)mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    if (out == MAP_FAILED) {
        ::close(fd);
        free(nindex);
        put_err(err, errlen, "cannot map the output file");
        return MISFIT_E_IO;
    }
    int32_t n32 = (int32_t)n, sentinel = -1;
    memcpy(out, &n32, 4);
    memcpy(out + 4, &sentinel, 4);
    memcpy(out + 8, &m, 8);
    memcpy(out + 16, nindex, ((size_t)n + 1) * 8);
    int32_t* nlist = (int32_t*)(out + hdr);
    // pass 2: entries in file order (nindex[u] walks forward as the cursor)
    fseek(r.f, data_pos, SEEK_SET);
    while ((len = r.next()) >= 0) {
        if (r.buf[0] == '%') continue;
        if (!two_fields(r.buf, a, b, nullptr)) continue;
        if (a <= n) nlist[nindex[a - 1]++] = (int32_t)(b - 1);
    }
    int rc = MISFIT_OK;
    if (msync(out, total, MS_SYNC) != 0) rc = MISFIT_E_IO;
    munmap(out, total);
    if (fsync(fd) != 0) rc = MISFIT_E_IO;
    ::close(fd);
    free(nindex);
    if (rc) put_err(err, errlen, "flush of the output file failed");
    if (n_out) *n_out = n32;
    if (m_out) *m_out = m;
    return rc;
}

// ---------------------------------------------------------------- validator
extern "C" int misfit_egr_check(const char* path, misfit_egr_report* rep, char* err, int errlen) {
    if (!path || !rep) return MISFIT_E_ARG;
    misfit_egr g;
    int rc = misfit_egr_open(path, &g, err, errlen);
    if (rc) return rc;
    memset(rep, 0, sizeof(*rep));
    rep->n = g.n;
    rep->m = g.m;
    rep->wide = g.wide;
    const int32_t n = g.n;
    auto idx = [&](int64_t i) -> int64_t {
        return g.wide ? ((const int64_t*)g.nindex)[i] : (int64_t)((const int32_t*)g.nindex)[i];
    };
    bool offsets_ok = idx(0) == 0 && idx(n) == g.m;
    if (!offsets_ok) rep->bad_offsets++;
    for (int32_t u = 0; u < n; u++)
        if (idx(u + 1) < idx(u)) rep->bad_offsets++;
    if (rep->bad_offsets) {  // rows cannot be walked safely
        rep->asymmetric = -1;
        misfit_egr_close(&g);
        return MISFIT_OK;
    }
    int64_t oor = 0, loops = 0, dups = 0, unsorted = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : oor, loops, dups, unsorted)
    for (int32_t u = 0; u < n; u++) {
        int64_t s = idx(u), e = idx(u + 1);
        bool sorted = true;
        for (int64_t p = s; p < e; p++) {
            int32_t v = g.nlist[p];
            if (v < 0 || v >= n) oor++;
            if (v == u) loops++;
            if (p > s && g.nlist[p - 1] >= v) sorted = false;
        }
        if (!sorted) {
            unsorted++;
            std::vector<int32_t> row(g.nlist + s, g.nlist + e);
            std::sort(row.begin(), row.end());
            for (size_t k = 1; k < row.size(); k++) dups += row[k] == row[k - 1];
        }
    }
    rep->out_of_range = oor;
    rep->self_loops = loops;
    rep->duplicates = dups;
    rep->unsorted_rows = unsorted;
    if (oor || unsorted) {
        rep->asymmetric = -1;  // symmetry is checked on ascending rows with in-range entries
    } else {
        int64_t asym = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : asym)
        for (int32_t u = 0; u < n; u++) {
            for (int64_t p = idx(u); p < idx(u + 1); p++) {
                int32_t v = g.nlist[p];
                const int32_t* b0 = g.nlist + idx(v);
                const int32_t* b1 = g.nlist + idx(v + 1);
                if (!std::binary_search(b0, b1, u)) asym++;
            }
        }
        rep->asymmetric = asym;
    }
    misfit_egr_close(&g);
    return MISFIT_OK;
}
