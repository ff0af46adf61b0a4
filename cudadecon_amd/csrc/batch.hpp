// Batched RL-BA over independent stacks of uneven shape (see batch.cpp).
#pragma once

#include "decon.hpp"

#include <memory>
#include <utility>
#include <vector>

namespace cdn {

// How Batch::run issues the iterations.
enum BatchMode : int {
  BATCH_SEQUENTIAL = 0,     // each stack's engine, one stack at a time
  BATCH_STREAMS = 1,        // each stack's engine, stacks round-robin over S streams
  BATCH_GRAPH = 2,          // mode 1 captured once into a hipGraph and replayed
  BATCH_GROUPED = 3,        // one launch per pass and transform length, spanning stacks
  BATCH_GROUPED_GRAPH = 4,  // mode 3 captured once into a hipGraph and replayed
};

// A captured schedule and what it baked in.
struct CapturedRun {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int iters = -1;
  std::vector<const void*> addr;                 // device addresses the launches use
  std::vector<std::pair<float*, float*>> post;   // each engine's X_k / X_{k-1} after a replay
  std::pair<float*, float*> arena_post{nullptr, nullptr};
  void reset();
  ~CapturedRun() { reset(); }
};

struct Batch {
  // one engine per stack (input order); its volumes and spectrum are views
  // of the arenas below, its OTF is table[table_of[i]] at its own k-scales
  std::vector<std::unique_ptr<RLEngine>> engines;
  std::vector<OtfTable> tables;
  std::vector<hipStream_t> streams;  // S working streams + 1 root
  std::vector<hipEvent_t> events;

  // arenas: every stack's raw, X ping-pong, Y, G (FP32) and spectrum, stacks
  // ordered by x length (then input order) so that the rows of all stacks of
  // one x length are contiguous
  DevBuf a_raw, a_xa, a_xb, a_y, a_g, a_spec;
  std::vector<int> order;        // arena position -> stack
  std::vector<int> pos;          // stack -> arena position
  std::vector<long> voff, soff;  // per stack: offset in the real arenas (floats), the spectrum arena (complex)
  float* cur = nullptr;          // arena X_k after the last grouped run
  float* prev = nullptr;

  // grouped schedule: one launch per pass and length (z filter: and table kind)
  struct XGroup {
    int nx = 0, first = 0, count = 0;  // arena positions first .. first + count - 1
    long rows = 0, voff = 0, soff = 0, pitch = 0, part0 = 0, nb = 0, rpb = 0;
    const float2* tw = nullptr;
    DevBuf row_starts;
  };
  struct ColGroup {
    int len = 0, nseg = 0;
    bool ra = false;  // z groups: every table rotationally averaged
    long total = 0;
    const float2* tw = nullptr;
    DevBuf segs;
  };
  // the grouped schedule of one partition: the stacks at arena positions
  // p0 .. p1 - 1, run on a stream of their own
  struct Partition {
    int p0 = 0, p1 = 0;
    std::vector<XGroup> xg;
    std::vector<ColGroup> yg, zg;
  };
  std::vector<Partition> parts;
  DevBuf g_part, g_lam, g_ranges;
  int lam_iters = 0;       // iterations whose factors g_lam holds (last grouped run)
  long grid_cap = 0;       // > 0: persistent column-pass grids of at most this many workgroups
  int npart = 1;           // partitions of the grouped schedule (concurrent, one stream each)
  int built_npart = 0;     // partitions the current tables were built for
  int last_mode = -1;

  CapturedRun cap_streams, cap_grouped;

  ~Batch();
  // dims_zyx: 3 ints per stack; kscales: 3 floats per stack; table_of: the
  // OTF table of each stack
  void setup(int n, const int* dims_zyx, std::vector<OtfTable> tabs, const int* table_of, const float* kscales,
             int nstreams);
  // iters iterations of every stack from its staged stack; returns device ms
  double run(int iters, int mode);
  // the factor of stack i at each iteration of the last run (0 for k < 2)
  std::vector<double> lambdas(int i, int iters) const;
  // kernel launches per iteration in a grouped run
  int grouped_launches() const;
  // partitions of the grouped schedule (1 <= n <= working streams); the
  // schedule is rebuilt at once
  void set_partitions(int n);

 private:
  void build_groups();
  void enqueue_engine(hipStream_t st, int i, int iters);
  void enqueue_streams(int iters);
  void enqueue_grouped(hipStream_t st, int iters);
  void enqueue_partition(const Partition& P, hipStream_t st, int iters, hipEvent_t mark);
  void capture(CapturedRun& c, int iters, bool grouped);
  std::vector<const void*> addresses(bool grouped) const;
};

}  // namespace cdn
