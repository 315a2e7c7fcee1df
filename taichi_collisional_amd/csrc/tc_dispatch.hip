// Routes the order-dispatched launches to the order shard that compiled them
// (tc_m2l.hip, tc_upward.hip and tc_downward.hip are built once per shard;
// shard k holds the orders p with p % TC_NSHARDS == k, see tc_kernels.h).
#include "tc_kernels.h"

namespace tcfmm {

#define TC_SHARD_API                                                                          \
  int launch_p2m(int p, CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves,            \
                 hipStream_t s, int what);                                                    \
  int launch_m2m(int p, CellsDev c, BodiesDev b, int ls, int le, hipStream_t s, int what);    \
  int launch_m2m_ids(int p, CellsDev c, BodiesDev b, const int *ids, int nids, hipStream_t s); \
  int launch_m2l(int p, const M2LArgs &a, hipStream_t s);                                     \
  int launch_l2l(int p, CellsDev c, int ls, int le, hipStream_t s);                           \
  int launch_l2l_ids(int p, CellsDev c, const int *ids, int nids, hipStream_t s);             \
  int launch_l2p(int p, CellsDev c, BodiesDev b, ResultsDev r, const int *leaf_ids,           \
                 int nleaves, hipStream_t s);

static_assert(TC_NSHARDS == 3, "the routing below names three shards");
namespace sh0 {
TC_SHARD_API
}
namespace sh1 {
TC_SHARD_API
}
namespace sh2 {
TC_SHARD_API
}

#define TC_ROUTE(call)                  \
  if (p < 0) return -1;                 \
  switch (p % TC_NSHARDS) {             \
    case 0:                             \
      return sh0::call;                 \
    case 1:                             \
      return sh1::call;                 \
    default:                            \
      return sh2::call;                 \
  }

int launch_p2m(int p, CellsDev c, BodiesDev b, const int *leaf_ids, int nleaves, hipStream_t s,
               int what) {
  TC_ROUTE(launch_p2m(p, c, b, leaf_ids, nleaves, s, what))
}
int launch_m2m(int p, CellsDev c, BodiesDev b, int ls, int le, hipStream_t s, int what) {
  TC_ROUTE(launch_m2m(p, c, b, ls, le, s, what))
}
int launch_m2m_ids(int p, CellsDev c, BodiesDev b, const int *ids, int nids, hipStream_t s) {
  TC_ROUTE(launch_m2m_ids(p, c, b, ids, nids, s))
}
int launch_m2l(int p, const M2LArgs &a, hipStream_t s) { TC_ROUTE(launch_m2l(p, a, s)) }
int launch_l2l(int p, CellsDev c, int ls, int le, hipStream_t s) {
  TC_ROUTE(launch_l2l(p, c, ls, le, s))
}
int launch_l2l_ids(int p, CellsDev c, const int *ids, int nids, hipStream_t s) {
  TC_ROUTE(launch_l2l_ids(p, c, ids, nids, s))
}
int launch_l2p(int p, CellsDev c, BodiesDev b, ResultsDev r, const int *leaf_ids, int nleaves,
               hipStream_t s) {
  TC_ROUTE(launch_l2p(p, c, b, r, leaf_ids, nleaves, s))
}

}  // namespace tcfmm
