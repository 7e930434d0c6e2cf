// SCC decomposition of a device-resident graph that allocates no memory: the
// per-vertex state, lists, bitmaps and control block live in one file-scope
// device array (`scratch`) instead of a context's allocations.  For callers
// that cannot allocate (code that must not call hipMalloc: graph-capture
// preparation, tight serving loops, harnesses that forbid it).  The pipeline
// and its kernels are the ones scc_run_method2 runs on a context; only where
// their buffers live differs.
#include <cstring>
#include <mutex>

#include "common.h"
#include "engine.h"
#include "kernels.h"

namespace sccmi {
namespace {

// SCC_NOALLOC_SCRATCH_BYTES of device memory on every GPU the library runs on,
// held until the process ends.  In host code this name stands for the device
// array only as a symbol: its device address comes from hipGetSymbolAddress.
__device__ __align__(256) unsigned char scratch[SCC_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};  // device address of `scratch`, per device

// Carves 256-byte aligned pieces off the scratch array, in a fixed order.
struct Carve {
  unsigned char* base;
  uint64_t off = 0;
  template <typename T>
  T* take(uint64_t count) {
    T* p = reinterpret_cast<T*>(base + off);
    off += (count * sizeof(T) + 255) & ~uint64_t(255);
    return p;
  }
};

static_assert(sizeof(DevCtl) < (1u << 20), "control block");

// max(out-degree, in-degree) over the vertices, into *out (which holds 0).
__global__ __launch_bounds__(BLOCK) void k_max_degree(int32_t n, const int32_t* __restrict__ begin,
                                                      const int32_t* __restrict__ rbegin, int32_t* __restrict__ out) {
  int32_t mx = 0;
  for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < n; v += (int64_t)gridDim.x * BLOCK)
    mx = max(mx, max(begin[v + 1] - begin[v], rbegin[v + 1] - rbegin[v]));
  mx = wave_max(mx);
  if (lane_id() == 0 && mx > 0) atomicMax(out, mx);
}
__global__ void k_zero_i32(int32_t* p) {
  if (threadIdx.x == 0) *p = 0;
}

int scratch_of(int device, unsigned char** out, std::string* err) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return SCC_ERR_HIP;
    }
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[device];
  return SCC_OK;
}

}  // namespace
}  // namespace sccmi

using namespace sccmi;

extern "C" int scc_run_device_noalloc(int device, void* stream, int32_t n, int64_t m, const int32_t* d_begin,
                                      const int32_t* d_node_idx, const int32_t* d_r_begin,
                                      const int32_t* d_r_node_idx, int32_t* d_labels, int64_t* num_sccs,
                                      char* err, size_t err_len) {
  std::string t_err;
  // the message into the caller's buffer on the way out (never a heap block of ours)
  struct Report {
    std::string& e;
    char* buf;
    size_t len;
    ~Report() {
      if (buf && len) {
        const size_t k = std::min(len - 1, e.size());
        std::memcpy(buf, e.data(), k);
        buf[k] = 0;
      }
    }
  } report{t_err, err, err_len};
  if (n < 0 || m < 0 || (n > 0 && (!d_begin || !d_r_begin || !d_labels)) ||
      (m > 0 && (!d_node_idx || !d_r_node_idx))) {
    t_err = "bad argument";
    return SCC_ERR_ARG;
  }
  if (n > SCC_NOALLOC_MAX_N || m >= INT32_MAX) {
    t_err = "graph exceeds the no-alloc limits (n <= " + std::to_string(SCC_NOALLOC_MAX_N) +
            " vertices, m < 2^31 edges); use a context";
    return SCC_ERR_LIMIT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= MAX_DEVICES) {
    t_err = "bad device";
    return SCC_ERR_ARG;
  }
  if (hipSetDevice(device) != hipSuccess) {
    t_err = "hipSetDevice failed";
    return SCC_ERR_HIP;
  }
  if (num_sccs) *num_sccs = 0;
  if (n == 0) return SCC_OK;
  unsigned char* base = nullptr;
  int rc = scratch_of(device, &base, &t_err);
  if (rc) return rc;

  // a context whose buffers are pieces of the scratch array and whose graph is
  // the caller's; dev_alloc refuses to allocate for it
  scc_ctx c;
  c.device = device;
  c.stream = static_cast<hipStream_t>(stream);
  c.no_alloc = true;
  c.n = n;
  c.m = m;
  c.m_cap = m;
  c.begin = const_cast<int32_t*>(d_begin);
  c.adj = const_cast<int32_t*>(d_node_idx);
  c.rbegin = const_cast<int32_t*>(d_r_begin);
  c.radj = const_cast<int32_t*>(d_r_node_idx);
  const uint64_t N = (uint64_t)n + 1, words = ((uint64_t)n + 31) / 32 + 1;
  Carve k{base};
  c.dc = k.take<DevCtl>(1);
  c.ctr = c.dc->ctr;
  int32_t* d_maxdeg = k.take<int32_t>(1);
  int32_t** arrs[] = {&c.color,   &c.label,    &c.canon,    &c.W,        &c.live[0], &c.live[1], &c.q[0],
                      &c.q[1],    &c.cand[0],  &c.cand[1],  &c.huge,     &c.huge2,   &c.bq[0],   &c.bq[1],
                      &c.bcand[0], &c.bcand[1], &c.bhuge[0], &c.bhuge[1], &c.scratch_i32};
  for (auto p : arrs) *p = k.take<int32_t>(N);
  c.vf = k.take<uint32_t>(2 * words);
  c.vb = c.vf + ((uint64_t)n + 31) / 32;
  c.pkey = k.take<unsigned long long>(N);
  c.rep = k.take<int32_t>(3 * N);
  c.cw = k.take<unsigned long long>(N);
  c.ckey = k.take<unsigned long long>(N);
  if (k.off > (uint64_t)SCC_NOALLOC_SCRATCH_BYTES) {  // guarded by the static check below
    t_err = "scratch layout exceeds the array";
    return SCC_ERR_LIMIT;
  }
  DevCtl h_dc;  // host mirror of the control block (pageable; no pinned allocation)
  std::memset(&h_dc, 0, sizeof(h_dc));
  c.h_dc = &h_dc;
  c.h_ctr = h_dc.ctr;

  auto fail = [&](int r) {
    t_err = c.err.empty() ? "HIP error" : c.err;
    c.h_dc = nullptr;
    return r;
  };
  // every control field and completion counter from scratch: the array keeps
  // the last call's contents
  launch_ctl_init(c.stream, c.dc, 1 | 2 | 4 | 8 | 16 | 32 | 64, 0, 0, 0, 0.0);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, c.stream, d_maxdeg);
  hipLaunchKernelGGL(k_max_degree, dim3(grid_for(n, 1024)), dim3(BLOCK), 0, c.stream, n, d_begin, d_r_begin,
                     d_maxdeg);
  int32_t maxdeg = 0;
  if ((rc = hip_check(&c, hipMemcpyAsync(&maxdeg, d_maxdeg, 4, hipMemcpyDeviceToHost, c.stream), "max degree")) ||
      (rc = hip_check(&c, hipStreamSynchronize(c.stream), "max degree sync")))
    return fail(rc);
  c.maxdeg = maxdeg;
  if ((rc = reset_state(&c))) return fail(rc);
  scc_run_stats st;
  if ((rc = run_method2(&c, nullptr, &st))) return fail(rc);
  if ((rc = hip_check(&c, hipMemcpyAsync(d_labels, c.canon, 4ull * (uint64_t)n, hipMemcpyDeviceToDevice, c.stream),
                      "labels copy")) ||
      (rc = hip_check(&c, hipStreamSynchronize(c.stream), "labels sync")))
    return fail(rc);
  if (num_sccs) *num_sccs = st.num_sccs;
  c.h_dc = nullptr;
  return SCC_OK;
}

namespace sccmi {
static_assert((uint64_t)SCC_NOALLOC_MAX_N * 120ull + (2ull << 20) <= (uint64_t)SCC_NOALLOC_SCRATCH_BYTES,  // 112 B/vertex carved
              "no-alloc scratch array too small for SCC_NOALLOC_MAX_N");
}  // namespace sccmi
