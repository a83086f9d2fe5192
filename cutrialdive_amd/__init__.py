"""MI355X-native trial-factoring engine (drop-in for cutrialdive's GPU path).

The compute runs in libctd_hip.so (HIP kernels for gfx950 behind the C ABI
of include/ctd_hip.h); this package holds the ctypes binding and a Python
mirror of the reference's host interface.
"""
from .engine import (  # noqa: F401
    Engine,
    FactoringResults,
    NumSeqId,
    NumSeqSpec,
    PEANO_SEQ_ID,
    Server,
    TrialFactoringOptions,
    num_seq_id_from_string,
    spec_param,
    trial_factor,
)

__all__ = [
    "Engine",
    "FactoringResults",
    "NumSeqId",
    "NumSeqSpec",
    "PEANO_SEQ_ID",
    "Server",
    "TrialFactoringOptions",
    "num_seq_id_from_string",
    "spec_param",
    "trial_factor",
]
