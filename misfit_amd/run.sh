#!/usr/bin/env bash
# run.sh -- build (if needed) and run one of the two MISFIT engines on MI355X.
# Same flags as the reference's run.sh (run.sh:1-127):
#
#   misfit_amd/run.sh --variant=misfit|misfit_async --graph=<graph.egr> \
#       --mis=<initial_mis.txt> --batch=<batch_folder> [--num-batches=N] [--gpu=ID]
#   misfit_amd/run.sh --variant=misfit_async --sample
#
# --num-batches defaults to the number of *.mtx files in the batch folder;
# --gpu defaults to 0.  --sample uses the bundled sample (tests/golden/sample);
# its graph is copied to a scratch directory first, so the end-of-run
# compaction rewrites the copy and never the bundled file.
#
# Environment: MISFIT_SKIP_COMPACTION=1 (leave the input .egr untouched),
# MISFIT_RULE=reference (literal rule, no completion), MISFIT_MODE=plain|queued,
# MISFIT_CHECK=0 (no check after every batch), HIPCC / MISFIT_ARCH (build:
# default /opt/rocm/bin/hipcc, gfx950).
set -euo pipefail

HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(dirname "$HERE")"

usage() {
    cat >&2 <<EOF
Usage: $0 --variant=misfit|misfit_async --graph=<graph.egr> --mis=<initial_mis.txt> --batch=<batch_folder> [--num-batches=N] [--gpu=ID]
       $0 --variant=misfit|misfit_async --sample

  --variant       misfit (one batch at a time) or misfit_async (batches queued back to back)
  --graph         base graph, MISFIT .egr binary CSR (int32 or int64-sentinel form)
  --mis           initial MIS, one 1-indexed vertex id per line
  --batch         folder of batch files named 1.mtx, 2.mtx, ...
  --num-batches   how many batch files to replay (default: count of *.mtx files in --batch)
  --gpu           HIP device id (default: 0)
  --sample        the small bundled sample dataset instead of --graph/--mis/--batch

Environment overrides:
  HIPCC                     hipcc binary to build with (default: /opt/rocm/bin/hipcc)
  MISFIT_ARCH               offload architecture (default: gfx950)
  MISFIT_SKIP_COMPACTION=1  skip the end-of-run .egr compaction (input graph left untouched)
  MISFIT_RULE=reference     literal reference rule, without the maximality completion
  MISFIT_MODE=plain|queued  host-driven or device-driven rounds (default: queued)
  MISFIT_CHECK=0            skip the independence/maximality check after every batch
EOF
}

VARIANT=""
GRAPH=""
MIS=""
BATCH=""
NUM_BATCHES=""
GPU="0"
USE_SAMPLE=0

for arg in "$@"; do
    case "$arg" in
        --variant=*)     VARIANT="${arg#--variant=}" ;;
        --graph=*)       GRAPH="${arg#--graph=}" ;;
        --mis=*)         MIS="${arg#--mis=}" ;;
        --batch=*)       BATCH="${arg#--batch=}" ;;
        --num-batches=*) NUM_BATCHES="${arg#--num-batches=}" ;;
        --gpu=*)         GPU="${arg#--gpu=}" ;;
        --sample)        USE_SAMPLE=1 ;;
        -h|--help)       usage; exit 0 ;;
        *)
            echo "ERROR: unrecognized argument: $arg" >&2
            usage
            exit 1
            ;;
    esac
done

if [ -z "$VARIANT" ]; then
    echo "ERROR: --variant=misfit|misfit_async is required" >&2
    usage
    exit 1
fi
case "$VARIANT" in
    misfit|misfit_async) BIN="$HERE/bin/$VARIANT" ;;
    *)
        echo "ERROR: --variant must be 'misfit' or 'misfit_async', got '$VARIANT'" >&2
        exit 1
        ;;
esac

if [ "$USE_SAMPLE" = "1" ]; then
    SAMPLE_DIR="$ROOT/tests/golden/sample"
    SCRATCH="$(mktemp -d "${TMPDIR:-/tmp}/misfit_sample.XXXXXX")"
    cp "$SAMPLE_DIR/sample.egr" "$SCRATCH/sample.egr"
    GRAPH="${GRAPH:-$SCRATCH/sample.egr}"
    MIS="${MIS:-$SAMPLE_DIR/sample_mis.txt}"
    BATCH="${BATCH:-$SAMPLE_DIR/batches}"
fi

if [ -z "$GRAPH" ] || [ -z "$MIS" ] || [ -z "$BATCH" ]; then
    echo "ERROR: --graph, --mis, and --batch are all required (or pass --sample)" >&2
    usage
    exit 1
fi
if [ ! -f "$GRAPH" ]; then
    echo "ERROR: graph not found: $GRAPH" >&2
    exit 1
fi
if [ ! -f "$MIS" ]; then
    echo "ERROR: initial MIS file not found: $MIS" >&2
    exit 1
fi
if [ ! -d "$BATCH" ]; then
    echo "ERROR: batch folder not found: $BATCH" >&2
    exit 1
fi

if [ -z "$NUM_BATCHES" ]; then
    NUM_BATCHES="$(find "$BATCH" -maxdepth 1 -name '*.mtx' | wc -l | tr -d ' ')"
    echo "--num-batches not given; found $NUM_BATCHES batch file(s) in $BATCH"
fi
case "$NUM_BATCHES" in
    ''|*[!0-9]*) echo "ERROR: num-batches must be a positive integer, got '$NUM_BATCHES'" >&2; exit 1 ;;
esac
case "$GPU" in
    ''|*[!0-9]*) echo "ERROR: --gpu must be a non-negative integer, got '$GPU'" >&2; exit 1 ;;
esac

if [ -x "$BIN" ]; then
    echo "binary already built: $BIN (delete it to force a rebuild)"
else
    echo "building $VARIANT -> $BIN"
    make -C "$HERE/csrc" -j8 HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}" ARCH="${MISFIT_ARCH:-gfx950}" \
        "../bin/$VARIANT"
fi

echo "=== running $VARIANT: graph=$GRAPH mis=$MIS batch=$BATCH num_batches=$NUM_BATCHES gpu=$GPU ==="
exec "$BIN" "$GRAPH" "$MIS" "$BATCH" "$NUM_BATCHES" "$GPU"
