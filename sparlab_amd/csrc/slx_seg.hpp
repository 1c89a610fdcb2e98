// Segmented (batched) CG kernels (kernels_seg.hip): independent systems on the
// diagonal blocks of one matrix, each kernel launch covering all of them.
#pragma once

#include <hip/hip_runtime.h>

namespace slx {
namespace kseg {

// Row chunks that never cross a segment boundary (device arrays).
struct Chunks {
  int count = 0;
  const int* begin = nullptr;      // first row of chunk c
  const int* end = nullptr;        // one past its last row
  const int* seg = nullptr;        // its segment
  const int* seg_chunk = nullptr;  // chunks of segment s: [seg_chunk[s], seg_chunk[s+1])
};

// out[s] = a.b over segment s (partials: one per chunk)
void dot(const Chunks& ch, int nseg, const double* a, const double* b, double* partials, double* out, hipStream_t s);
// active segments: alpha_s = rz_s / pq_s, x += alpha p, r -= alpha q; rr[s] = r.r over segment s
void update(const Chunks& ch, int nseg, double* x, double* r, const double* p, const double* q, const double* rz,
            const double* pq, const int* active, double* partials, double* rr, hipStream_t s);
// p = z + beta_s p
void direction(const Chunks& ch, double* p, const double* z, const double* beta, hipStream_t s);
// beta_s = (rzn_s - rzx_s) / rz_s (rzx null: rzn_s / rz_s; 0 when rz_s = 0); then rz = rzn
void beta(int nseg, const double* rzn, const double* rzx, double* rz, double* b, hipStream_t s);
// flag |= 1 if some row has a column outside its segment
void check_blocks(int n, const int* ptr, const int* col, const int* row_seg, const int* seg_ptr, int* flag,
                  hipStream_t s);

}  // namespace kseg
}  // namespace slx
