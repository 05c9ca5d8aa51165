#!/usr/bin/env bash
# run.sh -- build (when needed) and start one of the two drop-in MISFIT
# binaries on an MI355X.  The command line is the reference's (run.sh:17-36:
# the same flags, the same defaults, exit status 1 on any usage error):
#
#   misfit_amd/run.sh --variant=misfit|misfit_async --graph=<graph.egr> \
#       --mis=<initial_mis.txt> --batch=<batch_folder> [--num-batches=N] [--gpu=ID]
#   misfit_amd/run.sh --variant=misfit_async --sample
#
# Every flag also takes its value as the next word (--gpu 1).  The sample
# run copies the bundled graph to a scratch folder, so the compaction at the
# end of the run rewrites the copy and never the bundled file.
set -euo pipefail

readonly PKG_DIR="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
readonly REPO_DIR="$(dirname "$PKG_DIR")"
readonly SAMPLE_DIR="$REPO_DIR/tests/golden/sample"

print_help() {
    cat >&2 <<EOF
usage: $0 --variant=misfit|misfit_async --graph=<graph.egr> --mis=<initial_mis.txt> --batch=<batch_folder> [--num-batches=N] [--gpu=ID]
       $0 --variant=misfit|misfit_async --sample

flags
  --variant=NAME      misfit: one batch at a time; misfit_async: batches queued back to back
  --graph=FILE        base graph in MISFIT's .egr binary CSR (int32 or int64-sentinel nindex)
  --mis=FILE          starting MIS, one 1-indexed vertex id per line
  --batch=DIR         folder of update batches named 1.mtx, 2.mtx, ...
  --num-batches=N     batches to replay (default: the number of .mtx files in DIR)
  --gpu=ID            HIP device (default 0)
  --sample            the bundled sample; fills in --graph, --mis and --batch

environment
  HIPCC, MISFIT_ARCH        compiler and offload target of the build (/opt/rocm/bin/hipcc, gfx950)
  MISFIT_SKIP_COMPACTION=1  leave the input .egr as it is at the end of the run
  MISFIT_RULE=reference     the literal rule only, without the maximality completion
  MISFIT_MODE=plain|queued  rounds driven by the host or by the device (default queued)
  MISFIT_CHECK=0            no independence/maximality check after each batch
EOF
}

# die MESSAGE [help]: the message on stderr, the help text too when asked, status 1
die() {
    echo "run.sh: $1" >&2
    [ "${2:-}" = help ] && print_help
    exit 1
}

declare -A opt=([gpu]=0)
sample=no

while [ $# -gt 0 ]; do
    word="$1"
    shift
    case "$word" in
        -h | --help) print_help; exit 0 ;;
        --sample) sample=yes; continue ;;
        --variant | --graph | --mis | --batch | --num-batches | --gpu)
            [ $# -gt 0 ] || die "$word needs a value" help
            key="${word#--}"
            opt[$key]="$1"
            shift
            ;;
        --variant=* | --graph=* | --mis=* | --batch=* | --num-batches=* | --gpu=*)
            key="${word%%=*}"
            key="${key#--}"
            opt[$key]="${word#*=}"
            ;;
        *) die "unrecognized argument: $word" help ;;
    esac
done

variant="${opt[variant]:-}"
[ -n "$variant" ] || die "no engine chosen: pass --variant=misfit|misfit_async" help
[ "$variant" = misfit ] || [ "$variant" = misfit_async ] ||
    die "--variant must be 'misfit' or 'misfit_async' (got '$variant')"

if [ "$sample" = yes ]; then
    scratch="$(mktemp -d "${TMPDIR:-/tmp}/misfit_sample.XXXXXX")"
    cp "$SAMPLE_DIR/sample.egr" "$scratch/sample.egr"
    : "${opt[graph]:=$scratch/sample.egr}"
    : "${opt[mis]:=$SAMPLE_DIR/sample_mis.txt}"
    : "${opt[batch]:=$SAMPLE_DIR/batches}"
fi

for key in graph mis batch; do
    [ -n "${opt[$key]:-}" ] || die "--graph, --mis and --batch must all be given (or --sample)" help
done
[ -f "${opt[graph]}" ] || die "graph not found: ${opt[graph]}"
[ -f "${opt[mis]}" ] || die "initial MIS not found: ${opt[mis]}"
[ -d "${opt[batch]}" ] || die "batch folder not found: ${opt[batch]}"

is_count() { [[ "$1" =~ ^[0-9]+$ ]]; }

if [ -z "${opt[num-batches]:-}" ]; then
    shopt -s nullglob
    files=("${opt[batch]}"/*.mtx)
    shopt -u nullglob
    opt[num-batches]=${#files[@]}
    echo "replaying all ${opt[num-batches]} batch file(s) of ${opt[batch]} (no --num-batches)"
fi
is_count "${opt[num-batches]}" || die "--num-batches takes a whole number, not '${opt[num-batches]}'"
is_count "${opt[gpu]}" || die "--gpu takes a device number, not '${opt[gpu]}'"

binary="$PKG_DIR/bin/$variant"
if [ ! -x "$binary" ]; then
    echo "building $binary"
    make -C "$PKG_DIR/csrc" -j8 HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}" ARCH="${MISFIT_ARCH:-gfx950}" "../bin/$variant"
fi

echo "=== running $variant: graph=${opt[graph]} mis=${opt[mis]} batch=${opt[batch]} num_batches=${opt[num-batches]} gpu=${opt[gpu]} ==="
exec "$binary" "${opt[graph]}" "${opt[mis]}" "${opt[batch]}" "${opt[num-batches]}" "${opt[gpu]}"
