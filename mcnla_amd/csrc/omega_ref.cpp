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

namespace {

// One seed per OpenMP thread: seed_seq{seed} (col-block: first through the per-rank seed_seq).
std::vector<uint32_t> thread_seeds(uint32_t seed, bool col_block, int threads) {
  uint32_t s0 = seed;
  if (col_block) {  // one MPI rank: the first seed of seed_seq{seed}
    std::seed_seq seq{seed};
    uint32_t per_rank[1];
    seq.generate(per_rank, per_rank + 1);
    s0 = per_rank[0];
  }
  std::seed_seq seq2{s0};
  std::vector<uint32_t> per_thread(threads);
  seq2.generate(per_thread.begin(), per_thread.end());
  return per_thread;
}

}  // namespace

// Column-sampling indices: the same seed chain and chunking, raw mt19937 draws stored in a 32-bit
// index_t (values of 2^31 and above wrap to negative), abs(), then % ncol.
void column_indices_reference(uint32_t seed, bool col_block, int threads, int64_t len,
                              int64_t ncol, int64_t* out) {
  if (threads < 1) threads = 1;
  const std::vector<uint32_t> per_thread = thread_seeds(seed, col_block, threads);
  for (int t = 0; t < threads; ++t) {
    int64_t chunk = len / threads;
    const int64_t start = chunk * t;
    if (t == threads - 1) chunk = len - start;
    std::mt19937 gen(per_thread[t]);
    for (int64_t i = 0; i < chunk; ++i) {
      const int32_t v = static_cast<int32_t>(static_cast<uint32_t>(gen()));
      const int64_t a = v < 0 ? -static_cast<int64_t>(v) : v;
      out[start + i] = a % ncol;
    }
  }
}

void omega_reference(uint32_t seed, bool col_block, int threads, int64_t len, double* out) {
  if (threads < 1) threads = 1;
  const std::vector<uint32_t> per_thread = thread_seeds(seed, col_block, threads);
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
