// The reference's seeded Omega stream, produced on the host exactly as MCNLA 0.7.2 does without
// MKL, so a solve can use the same Omega as a reference run (the reference's own GPU sketcher
// also draws Omega on the host, col_block_gaussian_projection_sketcher_gpu.hpp:110):
//   seed chain    include/mcnla/core/random/streams/streams.hpp:118-150
//                 (std::seed_seq -> one seed per MPI rank (col-block only) -> std::seed_seq ->
//                 one seed per OpenMP thread -> std::mt19937 per thread)
//   chunking      include/mcnla/core/random/routine/gaussian.hpp:45-55 (thread i fills chunk i of
//                 len/threads values, the last thread the remainder)
//   distribution  include/mcnla/core/random/routine/gaussian_detail_nomkl.hh:37-40
//                 (std::normal_distribution<double>)
// This is the Omega *input* of a sketch, not a compute fallback: the sketch itself always runs
// on the GPU.
#include <cstdint>
#include <random>
#include <thread>
#include <vector>

namespace mcnla {

void omega_reference(uint32_t seed, bool col_block, int threads, int64_t len, double* out) {
  uint32_t s0 = seed;
  if (col_block) {  // one MPI rank: the first seed of seed_seq{seed}
    std::seed_seq seq{seed};
    uint32_t per_rank[1];
    seq.generate(per_rank, per_rank + 1);
    s0 = per_rank[0];
  }
  if (threads < 1) threads = 1;
  std::seed_seq seq2{s0};
  std::vector<uint32_t> per_thread(threads);
  seq2.generate(per_thread.begin(), per_thread.end());
  auto fill = [&](int t) {
    int64_t chunk = len / threads;
    const int64_t start = chunk * t;
    if (t == threads - 1) chunk = len - start;
    std::mt19937 gen(per_thread[t]);
    std::normal_distribution<double> dis(0.0, 1.0);
    for (int64_t i = 0; i < chunk; ++i) out[start + i] = dis(gen);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(fill, t);
  fill(0);
  for (auto& th : pool) th.join();
}

}  // namespace mcnla
