// Candidate search (reference: k_traverse, engine/engine.cuh:233-259): every leaf whose float box overlaps the
// probe's box (closed test), restricted to leaf >= probe for self-joins.
//
// Stackless: nodes are in depth-first preorder, so a lane walks forward through the node array, descending into
// the next node on overlap and jumping to the node's skip link otherwise. No per-lane stack (which would live in
// scratch memory on CDNA), one 16-byte + one 8-byte load per visited node, monotonically increasing addresses.
// Two passes with the same walk: count, then write at the scanned offsets (SoA pair buffers).
#include "engine.h"

namespace gsj {

namespace {

template <bool WRITE>
__global__ void __launch_bounds__(256) k_traverse(const float4* __restrict__ qbb, const uint32_t* __restrict__ qidx, uint32_t qbase, uint32_t nq,
                                                  const BNode* __restrict__ nodes, uint32_t nnodes, int upper, uint32_t* __restrict__ counts,
                                                  const uint64_t* __restrict__ offs, uint32_t* __restrict__ out_a, uint32_t* __restrict__ out_b) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const uint32_t q = qidx ? qidx[qbase + t] : qbase + t;
  const float4 b = qbb[q];
  uint64_t o = WRITE ? offs[t] : 0;
  uint32_t cnt = 0;
  uint32_t i = 0;
  // every step moves strictly forward (left child = i + 1, skip > i), so the walk ends within nnodes steps
  while (i < nnodes) {
    const float4 c = nodes[i].box;
    const bool ov = !(c.x > b.z || c.z < b.x || c.y > b.w || c.w < b.y);
    const uint32_t skip = (uint32_t)nodes[i].skip;
    uint32_t next;
    if (!ov) next = skip;
    else {
      const int32_t prim = nodes[i].prim;
      if (prim < 0) next = i + 1;
      else {
        if (!upper || (uint32_t)prim >= q) {
          if (WRITE) { out_a[o + cnt] = q; out_b[o + cnt] = (uint32_t)prim; }
          cnt++;
        }
        next = skip;
      }
    }
    if (next <= i) break;
    i = next;
  }
  if (!WRITE) counts[t] = cnt;
}

}  // namespace

void traverse(Context& c, const float4* qbb, const uint32_t* qidx, uint32_t qbase, uint32_t nq, const DBvh& t, bool upper, uint32_t* counts,
              const uint64_t* offs, uint32_t* out_a, uint32_t* out_b) {
  if (nq == 0) return;
  if (t.nnodes == 0) {
    if (!offs) GSJ_HIP(hipMemsetAsync(counts, 0, (size_t)nq * 4, c.stream));
    return;
  }
  if (!offs)
    hipLaunchKernelGGL(k_traverse<false>, dim3(div_up(nq, 256)), dim3(256), 0, c.stream, qbb, qidx, qbase, nq, t.nodes, t.nnodes, upper ? 1 : 0,
                       counts, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL(k_traverse<true>, dim3(div_up(nq, 256)), dim3(256), 0, c.stream, qbb, qidx, qbase, nq, t.nodes, t.nnodes, upper ? 1 : 0,
                       nullptr, offs, out_a, out_b);
  GSJ_LAUNCHED();
}

}  // namespace gsj
