// Host launchers of the SCC kernels (defined in kernels.hip).  Each takes the
// context's stream and device pointers; list lengths are host values.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sccmi {

// grid size for a list of n items (grid-stride beyond the cap)
inline unsigned grid_for(int64_t n, unsigned cap = 8192) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

void launch_fill_i32(hipStream_t st, int32_t* p, int32_t v, int64_t n);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, int64_t n);
void launch_iota(hipStream_t st, int32_t* p, int32_t n);
void launch_trim1(hipStream_t st, const int32_t* begin, const int32_t* adj, const int32_t* rbegin,
                  const int32_t* radj, int32_t* color, int32_t* label, const int32_t* list_in,
                  int64_t n_in, int32_t* list_out, unsigned long long* ctr);
void launch_trim2(hipStream_t st, const int32_t* begin, const int32_t* adj, const int32_t* rbegin,
                  const int32_t* radj, int32_t* color, int32_t* label, const int32_t* list_in,
                  int64_t n_in, unsigned long long* ctr);
void launch_reset_slots(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n,
                        const int32_t* color, unsigned long long* pkey, int32_t* rep);
void launch_pivot_key(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n,
                      const int32_t* color, const int32_t* begin, const int32_t* rbegin, int policy,
                      const uint32_t* scores, unsigned long long* pkey);
void launch_force_pivot(hipStream_t st, int32_t p, int32_t n, const int32_t* color,
                        unsigned long long* pkey);
void launch_compact_live(hipStream_t st, const int32_t* list, int64_t n_in, const int32_t* color,
                         int32_t* list_out, unsigned long long* ctr);
void launch_seed(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n, const int32_t* color,
                 const unsigned long long* pkey, uint32_t* V, int32_t* q_out,
                 unsigned long long* ctr);
void launch_td(hipStream_t st, const int32_t* off, const int32_t* nb, const int32_t* color,
               uint32_t* V, const int32_t* q_in, int64_t n_qin, int32_t* q_out, int32_t* huge,
               unsigned long long* ctr);
void launch_td_huge(hipStream_t st, const int32_t* off, const int32_t* nb, const int32_t* color,
                    uint32_t* V, const int32_t* huge, int64_t n_huge, int32_t* q_out,
                    unsigned long long* ctr);
void launch_bu(hipStream_t st, bool multi, const int32_t* roff, const int32_t* rnb,
               const int32_t* off, const int32_t* color, uint32_t* V, const int32_t* c_in,
               int64_t n_cin, int32_t* c_out, int32_t* q_out, unsigned long long* ctr);
void launch_classify(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n, int32_t* color,
                     int32_t* label, const uint32_t* vf, const uint32_t* vb,
                     const unsigned long long* pkey, int32_t* rep, int32_t* list_out,
                     unsigned long long* ctr);
void launch_recolor(hipStream_t st, const int32_t* list, int64_t n_in, int32_t n, int32_t* color,
                    const uint32_t* vf, const uint32_t* vb, const int32_t* rep);
void launch_wcc_init(hipStream_t st, const int32_t* list, int64_t n_in, const int32_t* begin,
                     int32_t* W);
void launch_wcc_iter(hipStream_t st, const int32_t* list, int64_t n_in, const int32_t* begin,
                     const int32_t* adj, const int32_t* color, int32_t* W, unsigned long long* ctr);
void launch_wcc_color(hipStream_t st, const int32_t* list, int64_t n_in, int32_t* color,
                      const int32_t* W, unsigned long long* ctr);
void launch_canon(hipStream_t st, int32_t n, const int32_t* label, int32_t* mn, int32_t* out,
                  unsigned long long* ctr);

}  // namespace sccmi
