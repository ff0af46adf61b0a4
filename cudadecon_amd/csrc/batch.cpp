// Batched RL-BA over many independent stacks of uneven shape on one GPU.
//
// Every stack's volumes and spectrum live in shared arenas, stacks ordered by
// x length, and each stack also has an engine (its own FFT plan, OTF table at
// its own k-scales, dot-product partials and factors) whose buffers are views
// of its slices.  Five ways to run n iterations of every stack:
//   0  each engine, one stack at a time on one stream (the reference's usage);
//   1  each engine, stacks round-robin over S streams;
//   2  schedule 1 captured once into a hipGraph (fork/join) and replayed;
//   3  grouped: every pass is one launch per distinct transform length,
//      spanning all stacks of that length whatever their other sizes.  The x
//      passes see the contiguous rows of all stacks of one x length (a stack
//      table gives each row pair its factor); the y passes and the z filter
//      take a table of per-stack segments (spectrum, column geometry, OTF) and
//      find each workgroup's stack by a search over it; the factors of all
//      stacks come from two launches over per-stack ranges of the partials.
//      About 3 nx + 4 ny + 2 nz + 2 launches per iteration (the distinct
//      lengths), instead of 9 or 10 per stack;
//   4  schedule 3 captured once into a hipGraph and replayed.
// The grouped schedule can run as P partitions: the stacks, in arena order,
// cut into P ranges of about equal voxel count, each range grouped on its
// own and run on a stream of its own, so that different passes of
// different partitions overlap (stacks at different passes fill HBM
// better than all stacks in one pass at a time).
// Every mode gives every stack bitwise the result of a one-stack engine: the
// arithmetic per column, row and partial block is the same, and each stack's
// partials are summed over the same blocks in the same order.
// The iteration issues no host synchronisation and no memset, so it can be
// captured; a stack restarts from its staged stack (X_0 = R) at each run.
#include "batch.hpp"

#include "rl_kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <map>

namespace cdn {

void CapturedRun::reset() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  iters = -1;
  addr.clear();
  post.clear();
}

Batch::~Batch() {
  cap_streams.reset();
  cap_grouped.reset();
  engines.clear();  // views of the arenas: release before the arenas
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  for (hipStream_t s : streams) (void)hipStreamDestroy(s);
}

void Batch::setup(int n, const int* dims_zyx, std::vector<OtfTable> tabs, const int* table_of, const float* kscales,
                  int nstreams) {
  if (n <= 0) fail("batch: no stacks");
  if (tabs.empty()) fail("batch: no OTF tables");
  for (int i = 0; i < n; ++i) {
    if (dims_zyx[3 * i] <= 0 || dims_zyx[3 * i + 1] <= 0 || dims_zyx[3 * i + 2] <= 0)
      fail("batch: stack " + std::to_string(i) + " has a non-positive size");
    if (table_of[i] < 0 || table_of[i] >= (int)tabs.size())
      fail("batch: stack " + std::to_string(i) + " names OTF table " + std::to_string(table_of[i]) + " of " +
           std::to_string(tabs.size()));
  }
  tables = std::move(tabs);
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
  // arena layout: stacks by x length, then input order; every stack of one
  // x length back to back (their rows contiguous), groups 256-byte aligned
  order.resize(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return dims_zyx[3 * a + 2] < dims_zyx[3 * b + 2]; });
  pos.assign(n, 0);
  voff.assign(n, 0);
  soff.assign(n, 0);
  long v = 0, c = 0;
  for (int p = 0; p < n; ++p) {
    const int i = order[p];
    const int nz = dims_zyx[3 * i], ny = dims_zyx[3 * i + 1], nx = dims_zyx[3 * i + 2];
    if (p > 0 && nx != dims_zyx[3 * order[p - 1] + 2]) {
      v = (v + 63) / 64 * 64;
      c = (c + 31) / 32 * 32;
    }
    pos[i] = p;
    voff[i] = v;
    soff[i] = c;
    v += (long)nx * ny * nz;
    c += (long)nz * ny * Fft3::pitch_of(nx);
  }
  a_raw.ensure(sizeof(float) * (size_t)v);
  a_xa.ensure(sizeof(float) * (size_t)v);
  a_xb.ensure(sizeof(float) * (size_t)v);
  a_y.ensure(sizeof(float) * (size_t)v);
  a_g.ensure(sizeof(float) * (size_t)v);
  a_spec.ensure(sizeof(float2) * (size_t)c);
  engines.resize(n);
  for (int i = 0; i < n; ++i) {
    const EngineViews ev{a_raw.as<float>() + voff[i], a_xa.as<float>() + voff[i], a_xb.as<float>() + voff[i],
                         a_y.as<float>() + voff[i],   a_g.as<float>() + voff[i],  a_spec.as<float2>() + soff[i]};
    engines[i] = std::make_unique<RLEngine>();
    engines[i]->setup(dims_zyx[3 * i + 2], dims_zyx[3 * i + 1], dims_zyx[3 * i], tables[table_of[i]],
                      kscales + 3 * i, streams[0], 1, &ev);
  }
  cur = a_xa.as<float>();
  prev = a_xb.as<float>();
  build_groups();
}

void Batch::build_groups() {
  const int n = (int)engines.size();
  parts.clear();
  // partitions: contiguous ranges of arena positions of about equal voxels
  const int P = std::max(1, std::min(npart, n));
  std::vector<int> cut{0};
  {
    long total = 0, acc = 0;
    for (int p = 0; p < n; ++p) total += engines[order[p]]->fft.nvox;
    for (int p = 0; p < n && (int)cut.size() < P; ++p) {
      acc += engines[order[p]]->fft.nvox;
      if (acc * P >= total * (long)cut.size() && p + 1 < n) cut.push_back(p + 1);
    }
    cut.push_back(n);
  }
  std::vector<int> ranges(2 * (size_t)n);
  long part_total = 0;
  for (size_t q = 0; q + 1 < cut.size(); ++q) {
    Partition P1;
    P1.p0 = cut[q];
    P1.p1 = cut[q + 1];
    // x passes: runs of stacks with one x length; a run ends after a stack
    // whose rows would put a row pair or an update workgroup across the
    // next stack (ny nz odd, or not a multiple of the workgroup's rows)
    for (int p = P1.p0; p < P1.p1;) {
      const RLEngine& e0 = *engines[order[p]];
      XGroup g;
      g.nx = e0.fft.nx;
      g.first = p;
      g.voff = voff[order[p]];
      g.soff = soff[order[p]];
      g.pitch = e0.fft.pitch;
      g.tw = e0.fft.twx.as<float2>();
      g.rpb = x_update_rows_per_block(g.nx);
      std::vector<long> starts;
      while (p < P1.p1 && engines[order[p]]->fft.nx == g.nx) {
        const long r = (long)engines[order[p]]->fft.ny * engines[order[p]]->fft.nz;
        starts.push_back(g.rows);
        ranges[2 * (size_t)p] = (int)(part_total + g.rows / g.rpb);
        ranges[2 * (size_t)p + 1] = (int)((r + g.rpb - 1) / g.rpb);
        g.rows += r;
        ++p;
        if (r % 2 != 0 || r % g.rpb != 0) break;
      }
      g.count = p - g.first;
      g.nb = x_update_blocks(g.nx, g.rows);
      g.part0 = part_total;
      part_total += g.nb;
      g.row_starts.ensure(sizeof(long) * starts.size());
      CDN_HIP(hipMemcpy(g.row_starts.p, starts.data(), sizeof(long) * starts.size(), hipMemcpyHostToDevice));
      P1.xg.push_back(std::move(g));
    }
    // y passes by y length, z filter by z length
    // z groups by length and table kind (rotationally averaged or 3-D), so
    // that each stack runs the z-filter kind of its one-stack engine
    using ZKey = std::pair<int, bool>;
    std::map<int, std::vector<ColSeg>> ys;
    std::map<ZKey, std::vector<ZSeg>> zs;
    std::map<int, const float2*> ytw;
    std::map<ZKey, const float2*> ztw;
    std::map<int, long> ytot;
    std::map<ZKey, long> ztot;
    for (int p = P1.p0; p < P1.p1; ++p) {
      const int i = order[p];
      const RLEngine& e = *engines[i];
      const Fft3& F = e.fft;
      float2* S = a_spec.as<float2>() + soff[i];
      {
        const long L = col_tile_cols(F.ny), nt = (F.nxh + L - 1) / L;
        ColSeg sg{S, F.nxh, F.pitch, (long)F.ny * F.pitch, nt, ytot[F.ny]};
        ys[F.ny].push_back(sg);
        ytot[F.ny] += (long)F.nz * nt;
        ytw[F.ny] = F.twy.as<float2>();
      }
      {
        const long L = col_tile_cols(F.nz), nt = (F.nxh + L - 1) / L;
        const OtfParams op = e.otf.params(e.table, 1.0 / (double)F.nvox);
        const ZKey key{F.nz, op.ra != 0};
        ZSeg sg{S, F.nxh, (long)F.ny * F.pitch, F.pitch, nt, F.ny, ztot[key], op};
        zs[key].push_back(sg);
        ztot[key] += (long)F.ny * nt;
        ztw[key] = F.twz.as<float2>();
      }
    }
    for (auto& kv : ys) {
      ColGroup g;
      g.len = kv.first;
      g.nseg = (int)kv.second.size();
      g.total = ytot[kv.first];
      g.tw = ytw[kv.first];
      g.segs.ensure(sizeof(ColSeg) * kv.second.size());
      CDN_HIP(hipMemcpy(g.segs.p, kv.second.data(), sizeof(ColSeg) * kv.second.size(), hipMemcpyHostToDevice));
      fft_prepare_ragged(g.len);
      P1.yg.push_back(std::move(g));
    }
    for (auto& kv : zs) {
      ColGroup g;
      g.len = kv.first.first;
      g.ra = kv.first.second;
      g.nseg = (int)kv.second.size();
      g.total = ztot[kv.first];
      g.tw = ztw[kv.first];
      g.segs.ensure(sizeof(ZSeg) * kv.second.size());
      CDN_HIP(hipMemcpy(g.segs.p, kv.second.data(), sizeof(ZSeg) * kv.second.size(), hipMemcpyHostToDevice));
      fft_prepare_ragged(g.len);
      P1.zg.push_back(std::move(g));
    }
    for (const XGroup& g : P1.xg) fft_prepare_ragged(g.nx);
    parts.push_back(std::move(P1));
  }
  g_part.ensure(sizeof(double) * 2 * (size_t)std::max(part_total, 1L));
  g_ranges.ensure(sizeof(int) * ranges.size());
  CDN_HIP(hipMemcpy(g_ranges.p, ranges.data(), sizeof(int) * ranges.size(), hipMemcpyHostToDevice));
  built_npart = npart;
  cap_grouped.reset();
}

int Batch::grouped_launches() const {
  int n = 0;
  for (const Partition& P : parts) n += 3 * (int)P.xg.size() + 4 * (int)P.yg.size() + 2 * (int)P.zg.size() + 2;
  return n;
}

void Batch::set_partitions(int n) {
  if (n < 1 || n > (int)streams.size() - 1) fail("batch: partitions must be 1 .. " + std::to_string(streams.size() - 1));
  npart = n;
  if (built_npart != npart) build_groups();
}

void Batch::enqueue_engine(hipStream_t st, int i, int iters) {
  RLEngine& e = *engines[i];
  e.reset(st);
  e.iterate(iters, 0, false, st, nullptr);
}

void Batch::enqueue_streams(int iters) {
  const int S = (int)streams.size() - 1;
  hipStream_t root = streams[S];
  CDN_HIP(hipEventRecord(events[0], root));
  for (int s = 0; s < S; ++s) CDN_HIP(hipStreamWaitEvent(streams[s], events[0], 0));
  for (int i = 0; i < (int)engines.size(); ++i) enqueue_engine(streams[i % S], i, iters);
  for (int s = 0; s < S; ++s) {
    CDN_HIP(hipEventRecord(events[1 + s], streams[s]));
    CDN_HIP(hipStreamWaitEvent(root, events[1 + s], 0));
  }
}

// After every pass the grouped schedule records `mark` on its stream, an
// event nothing waits on: measured on the 48-stack workload, 54.5-55.3 ->
// 55.5-56.9 Gvoxel/s in both the stream and the graph form (one at a time,
// unchanged: 55.2 -> 55.7-56.0).  Why a marker between two large dependent
// launches helps was not established.
void Batch::enqueue_partition(const Partition& P, hipStream_t st, int iters, hipEvent_t mark) {
  const int n = (int)engines.size();
  const float eps = FLT_EPSILON;  // src/RL-Biggs-Andrews.cpp:274
  float* c = a_xa.as<float>();
  float* pv = a_xb.as<float>();
  float2* spec = a_spec.as<float2>();
  for (int k = 0; k < iters; ++k) {
    double* lamk = g_lam.as<double>() + 4 * (size_t)n * k;
    if (k >= 2)
      accel_factor_ragged(st, g_part.as<double>(), g_ranges.as<int>() + 2 * P.p0, P.p1 - P.p0, (double)eps,
                          lamk + 4 * P.p0);
    const int mode = k >= 2 ? XIN_PRED : XIN_PLAIN;
    for (const XGroup& g : P.xg)
      launch_x_r2c(g.nx, st, c + g.voff, pv + g.voff, a_y.as<float>() + g.voff, lamk + 4 * g.first, mode,
                   spec + g.soff, g.rows, g.pitch, g.tw, 0, g.row_starts.as<long>(), g.count);
    CDN_HIP(hipEventRecord(mark, st));
    for (int f = 0; f < 2; ++f) {
      for (const ColGroup& g : P.yg)
        launch_col_ragged(g.len, 0, st, g.segs.as<ColSeg>(), g.nseg, g.total, g.tw, grid_cap);
      CDN_HIP(hipEventRecord(mark, st));
      for (const ColGroup& g : P.zg)
        launch_z_ragged(g.len, st, g.segs.as<ZSeg>(), g.nseg, g.total, f, g.tw, grid_cap, g.ra);
      CDN_HIP(hipEventRecord(mark, st));
      for (const ColGroup& g : P.yg)
        launch_col_ragged(g.len, 1, st, g.segs.as<ColSeg>(), g.nseg, g.total, g.tw, grid_cap);
      CDN_HIP(hipEventRecord(mark, st));
      if (f == 0) {
        for (const XGroup& g : P.xg)
          launch_x_lrcore(g.nx, st, spec + g.soff, a_raw.as<float>() + g.voff, g.rows, g.pitch, g.tw, eps);
        CDN_HIP(hipEventRecord(mark, st));
      }
    }
    const float* yv = k >= 2 ? a_y.as<float>() : c;
    for (const XGroup& g : P.xg) {
      const int nb = launch_x_update(g.nx, st, spec + g.soff, yv + g.voff, pv + g.voff, a_g.as<float>() + g.voff,
                                     g_part.as<double>() + 2 * g.part0, k >= 1, g.rows, g.pitch, g.tw);
      if (nb != g.nb) fail("batch: the x-update launch shape changed under the grouped schedule");
    }
    std::swap(c, pv);
  }
}

void Batch::enqueue_grouped(hipStream_t st, int iters) {
  const int n = (int)engines.size();
  copy_f32(st, a_raw.as<float>(), a_xa.as<float>(), a_raw.bytes / sizeof(float));  // X_0 = R, every stack
  // events: [0] fork, [1 .. S] joins, [S + 1 + q] partition q's pass marker
  const int S = (int)streams.size() - 1;
  if (parts.size() == 1) {
    enqueue_partition(parts[0], st, iters, events[S + 1]);
  } else {
    // fork: partition q on working stream q; join on `st`
    CDN_HIP(hipEventRecord(events[0], st));
    for (size_t q = 0; q < parts.size(); ++q) {
      hipStream_t sq = streams[q % S];
      CDN_HIP(hipStreamWaitEvent(sq, events[0], 0));
      enqueue_partition(parts[q], sq, iters, events[S + 1 + q % S]);
    }
    for (int s = 0; s < std::min<int>(S, (int)parts.size()); ++s) {
      CDN_HIP(hipEventRecord(events[1 + s], streams[s]));
      CDN_HIP(hipStreamWaitEvent(st, events[1 + s], 0));
    }
  }
  // X_k ends in xb after an odd count
  float* c = iters % 2 ? a_xb.as<float>() : a_xa.as<float>();
  float* pv = iters % 2 ? a_xa.as<float>() : a_xb.as<float>();
  cur = c;
  prev = pv;
  for (int i = 0; i < n; ++i) {
    engines[i]->cur = c + voff[i];
    engines[i]->prev = pv + voff[i];
  }
}

std::vector<const void*> Batch::addresses(bool grouped) const {
  std::vector<const void*> a;
  for (auto& e : engines) {
    const auto d = e->device_addresses();
    a.insert(a.end(), d.begin(), d.end());
  }
  if (grouped) {
    for (const DevBuf* b : {&a_raw, &a_xa, &a_xb, &a_y, &a_g, &a_spec, &g_part, &g_lam, &g_ranges}) a.push_back(b->p);
    for (const Partition& P : parts) {
      for (const XGroup& g : P.xg) a.push_back(g.row_starts.p);
      for (const ColGroup& g : P.yg) a.push_back(g.segs.p);
      for (const ColGroup& g : P.zg) a.push_back(g.segs.p);
    }
    a.push_back(reinterpret_cast<const void*>(grid_cap));
  }
  return a;
}

void Batch::capture(CapturedRun& cr, int iters, bool grouped) {
  // a graph replays the addresses it was captured with: reuse it only while
  // none has moved (a longer mode 0/1 run regrows an engine's lam, say)
  std::vector<const void*> addr = addresses(grouped);
  if (cr.exec && cr.iters == iters && cr.addr == addr) return;
  cr.reset();
  const int S = (int)streams.size() - 1;
  hipStream_t root = streams[S];
  CDN_HIP(hipDeviceSynchronize());
  CDN_HIP(hipStreamBeginCapture(root, hipStreamCaptureModeThreadLocal));
  if (grouped)
    enqueue_grouped(root, iters);
  else
    enqueue_streams(iters);
  CDN_HIP(hipStreamEndCapture(root, &cr.graph));
  CDN_HIP(hipGraphInstantiate(&cr.exec, cr.graph, nullptr, nullptr, 0));
  cr.iters = iters;
  cr.addr = std::move(addr);
  // capturing ran the host bookkeeping of the schedule: the X_k / X_{k-1}
  // pointers are now where a replay leaves them
  for (auto& e : engines) cr.post.emplace_back(e->cur, e->prev);
  cr.arena_post = {cur, prev};
}

double Batch::run(int iters, int mode) {
  if (iters <= 0) fail("batch: iteration count must be positive");
  if (mode < BATCH_SEQUENTIAL || mode > BATCH_GROUPED_GRAPH) fail("batch: unknown mode " + std::to_string(mode));
  const int S = (int)streams.size() - 1;
  hipStream_t root = streams[S];
  const bool grouped = mode == BATCH_GROUPED || mode == BATCH_GROUPED_GRAPH;
  // everything is allocated before any capture
  for (auto& e : engines) {
    e->timing = false;
    e->prepare(iters);
  }
  g_lam.ensure(sizeof(double) * 4 * engines.size() * (size_t)iters);
  if (grouped) {
    bool stale = built_npart != npart;
    for (const Partition& P : parts)
      for (const XGroup& g : P.xg)
        if (x_update_rows_per_block(g.nx) != g.rpb) stale = true;  // the autotuner changed an x launch shape
    if (stale) build_groups();
  }
  if (mode == BATCH_GRAPH) capture(cap_streams, iters, false);
  if (mode == BATCH_GROUPED_GRAPH) capture(cap_grouped, iters, true);
  CDN_HIP(hipDeviceSynchronize());
  hipEvent_t t0, t1;
  CDN_HIP(hipEventCreate(&t0));
  CDN_HIP(hipEventCreate(&t1));
  CDN_HIP(hipEventRecord(t0, root));
  if (mode == BATCH_SEQUENTIAL) {
    for (int i = 0; i < (int)engines.size(); ++i) enqueue_engine(root, i, iters);
  } else if (mode == BATCH_STREAMS) {
    enqueue_streams(iters);
  } else if (mode == BATCH_GROUPED) {
    enqueue_grouped(root, iters);
  } else {
    CapturedRun& cr = mode == BATCH_GRAPH ? cap_streams : cap_grouped;
    CDN_HIP(hipGraphLaunch(cr.exec, root));
    for (size_t i = 0; i < engines.size(); ++i) {
      engines[i]->cur = cr.post[i].first;
      engines[i]->prev = cr.post[i].second;
    }
    if (mode == BATCH_GROUPED_GRAPH) {
      cur = cr.arena_post.first;
      prev = cr.arena_post.second;
    }
  }
  CDN_HIP(hipEventRecord(t1, root));
  CDN_HIP(hipEventSynchronize(t1));
  float ms = 0.f;
  CDN_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  last_mode = mode;
  lam_iters = iters;
  return ms;
}

std::vector<double> Batch::lambdas(int i, int iters) const {
  if (i < 0 || i >= (int)engines.size()) fail("batch: stack index out of range");
  if (last_mode < 0) fail("batch: no run yet");
  if (iters > lam_iters) fail("batch: the last run had fewer iterations");
  std::vector<double> out((size_t)iters, 0.0);
  const bool grouped = last_mode == BATCH_GROUPED || last_mode == BATCH_GROUPED_GRAPH;
  const size_t n = engines.size();
  std::vector<double> h;
  if (grouped) {
    h.resize(4 * n * (size_t)iters);
    CDN_HIP(hipMemcpy(h.data(), g_lam.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  } else {
    h.resize(4 * (size_t)iters);
    CDN_HIP(hipMemcpy(h.data(), engines[i]->lam.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  }
  for (int k = 2; k < iters; ++k) out[k] = grouped ? h[4 * (n * k + pos[i])] : h[4 * (size_t)k];
  return out;
}

}  // namespace cdn
