// trk_mc_mx.hip -- the event code of trk_events.inc compiled as the mixed-precision path (FP32 sampling, FP64 state, ledger and tallies).
#include "trk_mc.h"

namespace trk {
namespace mx {
#define TRK_MX 1
#include "trk_events.inc"
#undef TRK_MX

template __global__ void k_transport<true, false>(McDev din);
template __global__ void k_transport<false, false>(McDev din);
template __global__ void k_transport<true, true>(McDev din);
template __global__ void k_transport<false, true>(McDev din);
template __global__ void k_batch<true>(BatchCtl *ctl, int n_groups);
template __global__ void k_batch<false>(BatchCtl *ctl, int n_groups);
}  // namespace mx
}  // namespace trk
