// Batched RL-BA over independent stacks (see batch.cpp).
#pragma once

#include "decon.hpp"

#include <memory>
#include <vector>

namespace cdn {

struct Batch {
  std::vector<std::unique_ptr<RLEngine>> engines;
  std::vector<hipStream_t> streams;  // S working streams + 1 root
  std::vector<hipEvent_t> events;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_iters = -1;
  // what the captured graph baked in: every engine's device addresses, and
  // each engine's X_k / X_{k-1} pointers after a replay
  std::vector<std::vector<const void*>> graph_addr;
  std::vector<std::pair<float*, float*>> graph_post;

  ~Batch();
  // dims_zyx: 3 ints per stack; kscales: 3 floats per stack
  void setup(int n, const int* dims_zyx, const OtfTable& t, const float* kscales, int nstreams);
  // iters iterations of every stack from its staged stack; returns device ms
  double run(int iters, int mode);
  void capture(int iters);
  void enqueue(hipStream_t st, int i, int iters);
};

}  // namespace cdn
