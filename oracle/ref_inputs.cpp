// ref_inputs.cpp -- regenerates the synthetic inputs behind the reference's
// published kernel checksums, so those checksums (data held in the reference's
// results/throughput/*.json) can pin this oracle and the HIP kernels.
//
// TEST INFRASTRUCTURE ONLY (loaded by tests/ and bench.py's cpu_baseline leg).
//
// Recipe restated from core/bench_kernels.cu:58-72 (complex cells) and
// core/bench_kernels.cu:128-147 (real recursive-torontonian cells):
// std::mt19937_64 seeded 1234 + matrix_dim, uniform_real_distribution(-1, 1),
// one (re, im) draw per entry in row-major order, optionally symmetrised by
// copying the upper triangle onto the lower one. The distribution's output is
// library specific, so this file must be compiled by g++ against libstdc++,
// as the reference's host code was.

#include <cstddef>
#include <random>

namespace {
struct pair_d {
  double re, im;
};
// Two-argument constructor call: the argument evaluation order is the
// compiler's, exactly as in the reference's make_cuDoubleComplex(U(rng), U(rng)).
inline pair_d make_pair_d(double re, double im) { return pair_d{re, im}; }
}  // namespace

extern "C" {

// out: batch * dim * dim complex128 (interleaved re, im)
void or_gen_bench_complex(int dim, int batch, int symmetric, double scale, double* out) {
  std::mt19937_64 rng(1234 + dim);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  pair_d* h = reinterpret_cast<pair_d*>(out);
  const int D = dim;
  for (int b = 0; b < batch; ++b)
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) h[((size_t)b * D + i) * D + j] = make_pair_d(scale * U(rng), scale * U(rng));
  if (symmetric)
    for (int b = 0; b < batch; ++b)
      for (int i = 0; i < D; ++i)
        for (int j = i + 1; j < D; ++j) h[((size_t)b * D + j) * D + i] = h[((size_t)b * D + i) * D + j];
}

// out: batch * dim * dim doubles, symmetric
void or_gen_bench_real(int dim, int batch, double scale, double* out) {
  std::mt19937_64 rng(1234 + dim);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  const int D = dim;
  for (int b = 0; b < batch; ++b)
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) out[((size_t)b * D + i) * D + j] = scale * U(rng);
  for (int b = 0; b < batch; ++b)
    for (int i = 0; i < D; ++i)
      for (int j = i + 1; j < D; ++j) out[((size_t)b * D + j) * D + i] = out[((size_t)b * D + i) * D + j];
}

}  // extern "C"
