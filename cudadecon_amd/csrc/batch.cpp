// Batched RL-BA over many independent stacks of uneven shape on one GPU.
//
// Each stack has its own RL engine (plan, OTF, HBM buffers).  Three ways to
// run n iterations of every stack:
//   0  one stack at a time on one stream (the reference's usage pattern);
//   1  stacks spread round-robin over S streams, so small kernels of
//      different stacks overlap and fill the chip;
//   2  the same schedule captured once into a hipGraph (fork/join over the S
//      streams) and replayed: one graph launch per batch run.
// The iteration issues no host synchronisation and no memset, so it can be
// captured; a stack restarts from its staged stack (X_0 = R) at each run.
#include "batch.hpp"

#include <algorithm>

namespace cdn {

Batch::~Batch() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  for (hipStream_t s : streams) (void)hipStreamDestroy(s);
}

void Batch::setup(int n, const int* dims_zyx, const OtfTable& t, const float* kscales, int nstreams) {
  if (n <= 0) fail("batch: no stacks");
  nstreams = std::max(1, std::min(nstreams, 32));
  for (int i = 0; i < nstreams + 1; ++i) {
    hipStream_t s;
    CDN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    streams.push_back(s);
  }
  for (int i = 0; i < 2 * nstreams + 1; ++i) {
    hipEvent_t e;
    CDN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    events.push_back(e);
  }
  engines.resize(n);
  for (int i = 0; i < n; ++i) {
    engines[i] = std::make_unique<RLEngine>();
    engines[i]->setup(dims_zyx[3 * i + 2], dims_zyx[3 * i + 1], dims_zyx[3 * i], t, kscales + 3 * i, streams[0]);
  }
}

void Batch::enqueue(hipStream_t st, int i, int iters) {
  RLEngine& e = *engines[i];
  e.reset(st);
  e.iterate(iters, 0, false, st, nullptr);
}

void Batch::capture(int iters) {
  for (auto& e : engines) e->prepare(iters);  // no allocation while capturing
  // a graph replays the addresses it was captured with: reuse it only while
  // no engine buffer has moved (a longer mode 0/1 run regrows lam, say)
  std::vector<std::vector<const void*>> addr;
  for (auto& e : engines) addr.push_back(e->device_addresses());
  if (exec && graph_iters == iters && addr == graph_addr) return;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  graph_iters = -1;
  const int S = (int)streams.size() - 1;
  hipStream_t root = streams[S];
  CDN_HIP(hipDeviceSynchronize());
  CDN_HIP(hipStreamBeginCapture(root, hipStreamCaptureModeThreadLocal));
  CDN_HIP(hipEventRecord(events[0], root));
  for (int s = 0; s < S; ++s) CDN_HIP(hipStreamWaitEvent(streams[s], events[0], 0));
  for (int i = 0; i < (int)engines.size(); ++i) enqueue(streams[i % S], i, iters);
  for (int s = 0; s < S; ++s) {
    CDN_HIP(hipEventRecord(events[1 + s], streams[s]));
    CDN_HIP(hipStreamWaitEvent(root, events[1 + s], 0));
  }
  CDN_HIP(hipStreamEndCapture(root, &graph));
  CDN_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  graph_iters = iters;
  graph_addr = std::move(addr);
  // capturing ran each engine's host bookkeeping of the schedule: its X_k and
  // X_{k-1} pointers are now where a replay leaves them
  graph_post.clear();
  for (auto& e : engines) graph_post.emplace_back(e->cur, e->prev);
}

double Batch::run(int iters, int mode) {
  const int S = (int)streams.size() - 1;
  hipStream_t root = streams[S];
  for (auto& e : engines) e->timing = false;
  if (mode == 2) capture(iters);
  CDN_HIP(hipDeviceSynchronize());
  hipEvent_t t0, t1;
  CDN_HIP(hipEventCreate(&t0));
  CDN_HIP(hipEventCreate(&t1));
  CDN_HIP(hipEventRecord(t0, root));
  if (mode == 0) {
    for (int i = 0; i < (int)engines.size(); ++i) enqueue(root, i, iters);
  } else if (mode == 1) {
    CDN_HIP(hipEventRecord(events[0], root));
    for (int s = 0; s < S; ++s) CDN_HIP(hipStreamWaitEvent(streams[s], events[0], 0));
    for (int i = 0; i < (int)engines.size(); ++i) enqueue(streams[i % S], i, iters);
    for (int s = 0; s < S; ++s) {
      CDN_HIP(hipEventRecord(events[1 + s], streams[s]));
      CDN_HIP(hipStreamWaitEvent(root, events[1 + s], 0));
    }
  } else {
    CDN_HIP(hipGraphLaunch(exec, root));
    for (size_t i = 0; i < engines.size(); ++i) {
      engines[i]->cur = graph_post[i].first;
      engines[i]->prev = graph_post[i].second;
    }
  }
  CDN_HIP(hipEventRecord(t1, root));
  CDN_HIP(hipEventSynchronize(t1));
  float ms = 0.f;
  CDN_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return ms;
}

}  // namespace cdn
