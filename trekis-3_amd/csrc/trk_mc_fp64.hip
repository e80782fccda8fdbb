// trk_mc_fp64.hip -- the event code of trk_events.inc compiled as the deterministic FP64 path (replays the oracle bit for bit).
#include "trk_mc.h"

namespace trk {
namespace fp64 {
#define TRK_MX 0
#include "trk_events.inc"
#undef TRK_MX

template __global__ void k_transport<true, false>(McDev din);
template __global__ void k_transport<false, false>(McDev din);
template __global__ void k_transport<true, true>(McDev din);
template __global__ void k_transport<false, true>(McDev din);
template __global__ void k_batch<true>(BatchCtl *ctl, int n_groups);
template __global__ void k_batch<false>(BatchCtl *ctl, int n_groups);
}  // namespace fp64
}  // namespace trk
