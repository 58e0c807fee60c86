#!/bin/bash
# Run the narrow-pulse (default) or gradient simulation CLI on one GPU
# (run-scripts/run-simulation.sh of the reference, with rocm-smi for the GPU
# check and HIP_VISIBLE_DEVICES for the selection).
#   run-simulation.sh --gpu 0 --substrates DIR [simulation|gradient_simulation] [options]
GPU_DEVICE=0
SUBSTRATES_PATH=""
while [[ $# -gt 0 ]]; do
    case $1 in
        --gpu=*|--gpus=*) GPU_DEVICE="${1#*=}"; shift ;;
        --gpu|--gpus) GPU_DEVICE="$2"; shift 2 ;;
        --substrates=*) SUBSTRATES_PATH="${1#*=}"; shift ;;
        --substrates) SUBSTRATES_PATH="$2"; shift 2 ;;
        *) break ;;
    esac
done
COMMAND=simulation
if [[ "$1" == "simulation" || "$1" == "gradient_simulation" ]]; then COMMAND="$1"; shift; fi
if ! [[ "$GPU_DEVICE" =~ ^[0-9]+$ ]]; then
    echo "Error: '$GPU_DEVICE' is not a valid GPU number (one integer)"; exit 1
fi
if command -v rocm-smi &> /dev/null; then
    NGPU=$(rocm-smi --showid --csv 2>/dev/null | grep -c '^card')
    if [ "$NGPU" -gt 0 ] && [ "$GPU_DEVICE" -ge "$NGPU" ]; then
        echo "Error: GPU $GPU_DEVICE does not exist (valid: 0-$((NGPU-1)))"; exit 1
    fi
else
    echo "Warning: rocm-smi not found. Cannot validate the GPU number."
fi
if [ -z "$SUBSTRATES_PATH" ]; then echo "Error: --substrates is required"; exit 1; fi
echo "Using GPU: $GPU_DEVICE"
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
cd "$HERE/../.." && HIP_VISIBLE_DEVICES=$GPU_DEVICE exec python3 -m nozomi_amd.cli.main "$COMMAND" --substrates="$SUBSTRATES_PATH" "$@"
