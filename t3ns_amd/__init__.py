"""t3ns_amd: MI355X-native effective-Hamiltonian matvec, Davidson and
operator renormalisation for block-sparse three-legged tree tensor networks
(the hot path of the T3NS code), behind the C ABI in include/t3ns_amd.h."""
from ._lib import BF16, FP32, FP64, lib  # noqa: F401
from .heff import (Context, Heff, HeffBatch, ROps, Step, append_phys, clear_tune_cache,  # noqa: F401
                   davidson_device, davidson_refined, make_site_tensor, plan_block_dims,
                   plan_table, qr_step, split_site, update_branching, update_physical)

__all__ = ["Context", "Step", "Heff", "plan_table", "plan_block_dims", "FP64", "FP32", "BF16", "lib"]
