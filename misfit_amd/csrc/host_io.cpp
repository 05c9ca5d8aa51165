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

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>

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
