"""cudamica_amd -- MI355X-native AMICA (adaptive mixture ICA) EM engine.

The EM iteration runs in hand-written HIP kernels for gfx950 behind the C
ABI of include/amica_engine.h; `cudaamica -f <file>` (bin/) is the drop-in
command line program with the reference's parameter-file format.
"""
from .engine import (BATCH_POLICIES, CLI_PATH, LIB_PATH, PARAM_NAMES, PREC_BF16, PREC_FP8, PREC_FP32, PREC_FP32_PLAIN,
                     Batch,
                     Engine, EngineError,
                     cond, hbm_probe, ll_adjoint, ll_forward, ll_forward_noalloc, load_library, options, synth_data, synth_init,
                     synth_spec)
from .workloads import WORKLOADS, Workload, batch_workload

__all__ = ["Engine", "EngineError", "options", "load_library", "synth_data", "synth_init",
           "synth_spec", "cond", "PREC_FP32", "PREC_FP32_PLAIN", "PREC_BF16", "PREC_FP8", "LIB_PATH", "CLI_PATH", "WORKLOADS",
           "Workload", "hbm_probe", "ll_forward", "ll_forward_noalloc", "ll_adjoint", "PARAM_NAMES", "Batch", "BATCH_POLICIES", "batch_workload"]
