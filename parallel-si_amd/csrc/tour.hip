// Euler-tour path sums over a forest of links: the valuation engine of the
// reference's GPUList (gpu_list_ranking.cu:52-92 builds the tour of the sink
// tree, :96-277 ranks it by splitters, sublist walks, the reduced list and a
// broadcast), with one 64-bit weight per vertex where the reference carries a
// dense count vector per vertex.  For weights 1 the sum is the path length;
// for weights splitmix64(priority) it is the valuation digest (the multiset
// of priorities on the path, hashed additively).
//
// Tour: vertex v has an enter node 2v and an exit node 2v+1; the root (the
// sink) is vertex n with head 2n and tail 2n+1.  Each vertex pushes itself
// onto its parent's child list with one atomic exchange (child order is the
// exchange order; sums do not depend on it), except the root's children,
// which are many (every vertex whose link is -1) and would serialise on one
// word: their list is the compacted set {v : link[v] = -1} in id order,
// linked in one pass.  enter(v) leads to its first
// child's enter node or to its own exit node; exit(v) leads to the next
// sibling's enter node or to the parent's exit node.  The tour from the head
// visits exactly the vertices whose links reach the root; a link chain that
// closes a cycle (or ends at a link below -1) forms components the head never
// reaches.  Ranking the tour gives each node its position; +w at enter(v) and
// -w at exit(v), scanned in tour order, give at enter(v) the sum over v and its
// ancestors.
#include "common.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace psi {
namespace {

constexpr int kTB = 256;

__global__ void k_tour_first(int n, int* __restrict__ first, int* __restrict__ head)
{
    int v = blockIdx.x * kTB + threadIdx.x;
    if (v <= n) first[v] = -1;
    if (v == 0) *head = 2 * n;
}

__device__ __forceinline__ int tour_parent(const int* __restrict__ link, int n, int v)
{
    const int l = link[v];
    return l >= 0 ? l : (l == -1 ? n : -1);        // -1: a detached root
}

// Children of vertices (the root's are linked by k_tour_root).
__global__ void k_tour_push(int n, const int* __restrict__ link, int* __restrict__ first,
                            int* __restrict__ sib)
{
    int v = blockIdx.x * kTB + threadIdx.x;
    if (v >= n) return;
    const int p = tour_parent(link, n, v);
    if (p == n) return;
    sib[v] = p >= 0 ? atomicExch(&first[p], v) : -1;
}

// The root's children: kids[0 .. *nk) are the vertices with link -1.
__global__ void k_tour_root(int n, const int* __restrict__ kids, const int* __restrict__ nk,
                            int* __restrict__ first, int* __restrict__ sib)
{
    const int i = blockIdx.x * kTB + threadIdx.x;
    const int k = *nk;
    if (i >= k) return;
    sib[kids[i]] = i + 1 < k ? kids[i + 1] : -1;
    if (i == 0) first[n] = kids[0];
}

struct IsRootChild {
    __host__ __device__ bool operator()(const int& l) const { return l == -1; }
};

__global__ void k_tour_succ(int n, const int* __restrict__ link, const int* __restrict__ first,
                            const int* __restrict__ sib, int* __restrict__ succ)
{
    int v = blockIdx.x * kTB + threadIdx.x;
    if (v > n) return;
    succ[2 * v] = first[v] >= 0 ? 2 * first[v] : 2 * v + 1;
    if (v == n) {
        succ[2 * v + 1] = -1;                       // the tail
        return;
    }
    const int p = tour_parent(link, n, v);
    succ[2 * v + 1] = sib[v] >= 0 ? 2 * sib[v] : (p >= 0 ? 2 * p + 1 : -1);
}

__global__ void k_tour_scatter(int n, const int* __restrict__ rank, const uint64_t* __restrict__ w,
                               uint64_t* __restrict__ a)
{
    long long x = (long long)blockIdx.x * kTB + threadIdx.x;
    if (x >= 2LL * n + 2) return;
    const int r = rank[x];
    if (r < 0) return;
    const int v = (int)(x >> 1);
    const uint64_t wv = v < n ? w[v] : 0ull;
    a[r] = (x & 1) ? (0ull - wv) : wv;
}

__global__ void k_tour_gather(int n, const int* __restrict__ rank, const uint64_t* __restrict__ s,
                              uint64_t* __restrict__ sum, int* __restrict__ reach)
{
    int v = blockIdx.x * kTB + threadIdx.x;
    if (v >= n) return;
    const int r = rank[2 * v];
    sum[v] = r >= 0 ? s[r] : 0ull;
    if (reach) reach[v] = r >= 0;
}

template <class T> __global__ void k_tour_zero(T* p, long long n)
{
    long long i = (long long)blockIdx.x * kTB + threadIdx.x;
    if (i < n) p[i] = 0;
}

struct TourWork {
    void* buf = nullptr;
    size_t cap = 0;
};
TourWork g_tour;

} // namespace
} // namespace psi

using namespace psi;

extern "C" int psi_tour_path_sums(const int* d_link, int n, const uint64_t* d_w, uint64_t* d_sum,
                                  int* d_reach, int method)
{
    if (n < 0) return fail(PSI_ERR_INVALID, "psi_tour_path_sums: negative size");
    if (n == 0) return PSI_OK;
    if (n >= (1 << 30) - 1) return fail(PSI_ERR_INVALID, "psi_tour_path_sums: tour does not fit in int32");
    if ((method & 0xff) == PSI_LR_WYLLIE)
        return fail(PSI_ERR_INVALID, "psi_tour_path_sums: Wyllie ranks from every node without a "
                                     "predecessor; use HJ, WALK or AUTO");
    hipStream_t s = stream();
    const long long L = 2LL * n + 2;
    size_t scan_bytes = 0, sel_bytes = 0;
    PSI_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (const uint64_t*)nullptr,
                                             (uint64_t*)nullptr, (int)L, s));
    const hipcub::CountingInputIterator<int> ids(0);
    const hipcub::TransformInputIterator<bool, IsRootChild, const int*> is_kid(d_link, IsRootChild());
    PSI_HIP(hipcub::DeviceSelect::Flagged(nullptr, sel_bytes, ids, is_kid, (int*)nullptr,
                                          (int*)nullptr, n, s));
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_first = al((n + 1) * sizeof(int)), b_sib = al(n * sizeof(int));
    const size_t b_succ = al(L * sizeof(int)), b_rank = al(L * sizeof(int));
    const size_t b_a = al(L * sizeof(uint64_t)), b_s = al(L * sizeof(uint64_t));
    const size_t b_tmp = al(std::max(scan_bytes, sel_bytes));   // the select, then the scan
    const size_t need = b_first + b_sib + b_succ + b_rank + b_a + b_s + b_tmp + 4 * sizeof(int);
    if (int rc = ensure(&g_tour.buf, &g_tour.cap, need)) return rc;
    char* p = (char*)g_tour.buf;
    int* first = (int*)p; p += b_first;
    int* sib = (int*)p; p += b_sib;
    int* succ = (int*)p; p += b_succ;
    int* rank = (int*)p; p += b_rank;
    uint64_t* a = (uint64_t*)p; p += b_a;
    uint64_t* sc = (uint64_t*)p; p += b_s;
    void* tmp = p; p += b_tmp;
    int* head = (int*)p;
    int* nk = head + 1;
    int* kids = rank;                               // free until the ranking (n <= L ints)
    k_tour_first<<<nblocks(n + 1, kTB), kTB, 0, s>>>(n, first, head);
    k_tour_push<<<nblocks(n, kTB), kTB, 0, s>>>(n, d_link, first, sib);
    PSI_HIP(hipcub::DeviceSelect::Flagged(tmp, sel_bytes, ids, is_kid, kids, nk, n, s));
    k_tour_root<<<nblocks(n, kTB), kTB, 0, s>>>(n, kids, nk, first, sib);
    k_tour_succ<<<nblocks(n + 1, kTB), kTB, 0, s>>>(n, d_link, first, sib, succ);
    k_tour_zero<uint64_t><<<nblocks(L, kTB), kTB, 0, s>>>(a, L);
    PSI_LAUNCH_CHECK();
    if (int rc = psi_list_rank(succ, (int)L, head, 1, rank, (method & 0xff) | PSI_LR_CHECKED)) return rc;
    k_tour_scatter<<<nblocks(L, kTB), kTB, 0, s>>>(n, rank, d_w, a);
    PSI_LAUNCH_CHECK();
    PSI_HIP(hipcub::DeviceScan::InclusiveSum(tmp, scan_bytes, a, sc, (int)L, s));
    k_tour_gather<<<nblocks(n, kTB), kTB, 0, s>>>(n, rank, sc, d_sum, d_reach);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}
