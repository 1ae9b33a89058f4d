"""cdlm -- host side of the MI355X CDL-ADMM / PWM-scan library.

Mirrors the reference's Julia entry points for the hot path of CDLmotif.jl:
  CDL_setup, init_cdl, code_update!, dict_update!, update_best!, objective, cdl_optimize
  (src/SETUP/WRAPPER_setup.jl, src/SETUP/CDL_setup.jl, src/CDL/train.jl, src/WRAPPER/wrap.jl)
  scan_w_gpu!, overlapping_scan!, pfm2pwm, pwm_complement, best_score
  (src/SEARCH/scan.jl, src/SEARCH/helpers.jl, src/SEARCH/PWM_Touzet.jl)
All compute runs in libcdlm.so on a gfx950 device.
"""
from ._lib import CdlmError, CdlmOOM, load  # noqa: F401
from .device import Context, default_context  # noqa: F401
from .cdl import (CDL, Batch, footprint_bytes, pack_by_footprint, autotune, tuning_set, tuning_get, CDL_setup,  # noqa: F401
                  DNADataset, init_cdl, code_update, dict_update, update_best, objective, cdl_optimize)
from .serve import Server  # noqa: F401
from . import io  # noqa: F401
from .record import (ActivationRecord, record_training_result, get_spatial_info, topology,  # noqa: F401
                     overlap_pairs, get_pair_pfms)
from .scan import (pfm2pwm, pwm_complement, best_score, pack_pwms, scan_hits, scan_w_gpu,  # noqa: F401
                   overlapping_scan, onehot_to_bases, get_thresh, pvalue_last_ms)
from .evalue import fisher_pvalue, evalues, filter_using_evalue, sequences_with_hits  # noqa: F401
