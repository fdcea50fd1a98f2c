"""theeggman: MI355X (gfx950) engine for TheEggman.jl's subset-DP (`method=:dp`) hafnian path.

Python mirror of the reference's public interface for that path -- ``hafnian``,
``hafnian_repeated``, the plan (``_dp_plan``) and the device hook (``_haf_dp_backend``) with
its caches -- over the C ABI of ``libeggman_hip.so`` (include/eggman_hip.h).  Importing the
package loads the library; there is no CPU fallback.
"""
from ._lib import (ArgumentError, DimensionMismatch, EggmanError, F32, F64, C32, C64, I128,
                   lib as _load_lib)
from .plan import (DP_MAX, DP_HARD_MAX, UNROLL_MAX, HafnianDPPlan, PlanRef, _build_dp_plan, _dp_counts,
                   _dp_index_type, _dp_plan, _pair_bits, _pair_index, _pair_mask)
from .backend import (HIPBackend, _device_plan, _haf_dp_backend, _haf_dp_backend_grad,
                      _haf_dp_backend_mixed,
                      _haf_dp_device,
                      dp_batch_bytes, empty_gpu_cache, gpu_cache_bytes, grad_levels)
from .hafnian import (_gather_pairs_batch, default_backend, hafnian, hafnian_exact,
                      hafnian_grad, hafnian_mixed, hafnian_repeated, hafnian_repeated_grad,
                      reduction, rrule)

from .tuning import autotune, load_tuning, save_tuning, tune_get, tune_set

_load_lib()


def _load_default_tuning():
    """Launch choices measured on an MI355X ship beside the package; load them when a GPU is
    visible (every choice gives the same results, so this only affects speed)."""
    import os
    from .tuning import DEFAULT_TABLE
    if os.path.exists(DEFAULT_TABLE) and _load_lib().eggman_device_count() > 0:
        load_tuning(DEFAULT_TABLE)


_load_default_tuning()

__all__ = ["hafnian", "hafnian_repeated", "hafnian_exact", "hafnian_grad", "hafnian_mixed",
           "hafnian_repeated_grad", "rrule", "reduction", "HIPBackend",
           "gpu_cache_bytes", "empty_gpu_cache", "dp_batch_bytes", "DP_MAX", "DP_HARD_MAX",
           "HafnianDPPlan", "ArgumentError", "DimensionMismatch", "EggmanError"]
