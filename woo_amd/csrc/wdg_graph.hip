// Graph replay of the step loop and its autotuner.
//
// The engine list [Collider, ContactLoop, Leapfrog(reset)] runs as: the collider on the
// host (its dirty test reads one flag word, a rebuild sizes buffers), then K steps of contact
// loop and leapfrog captured once as a hipGraph and replayed. Before step j > 0 of a replay a
// one-thread kernel reads the dirty word the leapfrog of step j-1 set: when a particle left
// its Verlet skin it sets the stop word, and every captured kernel after it returns at once,
// so the replay ends after step j-1 and the host rebuilds before the next step, exactly as
// Scene::doOneStep would. The same kernel counts the executed steps and sets the step's
// trsfRenorm flag from the device-resident step number. Host bookkeeping (which state
// buffer is current, the parity of the dirty words) advances by the executed count.
//
// A graph is valid while the slot list, the buffers and the parameters it captured stay the
// same: it is captured again after a rebuild or any change of state layout (Ctx::gen), of
// the tuning, or of the parity of the state buffers. K is even so that a replay ends on the
// buffer parity it started with.
//
// The autotuner times candidate tunings (workgroups per contact group in the force kernel,
// steps per graph) on the live state: each trial is a run of real steps, timed by the host
// clock around whole replays, the best kept.
#include <chrono>

#include "wdg_internal.hpp"

namespace wdg {

// gstate words: [0] stop, [1] steps executed in the batch, [2] renorm flag of the step, [3..4]
// the batch's first step (int64), [5] trsfRenorm, [6] set until the batch's first step starts.
// A batch is several replays launched back to back: once a gate has stopped, every later kernel
// of the batch returns at once, so the replays after the stop cost only their launches.
// dirtyWord: the word the previous step's leapfrog sets (its parity is fixed at capture; K is even)
__global__ void k_gate(uint32_t* gs, const uint32_t* dflags, int dirtyWord) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (gs[0]) return;
  // the batch's first step follows the host's collider call, which has rebuilt if needed
  if (!gs[6] && dflags[dirtyWord]) {
    gs[0] = 1u;
    return;
  }
  gs[6] = 0u;
  const uint32_t ex = gs[1];
  gs[1] = ex + 1u;
  const int64_t step = (int64_t)(((uint64_t)gs[4] << 32) | gs[3]) + ex;
  const int64_t tr = (int64_t)gs[5];
  gs[2] = (tr > 0 && (step % tr) == 0) ? 1u : 0u;
}

// [Collider, ContactLoop, (IntraForce,) Leapfrog]
static bool has_intra(const wdg_program* prog) { return prog->nEngines == 4; }
static void check_program(const wdg_program* prog) {
  const int n = prog->nEngines;
  if ((n != 3 && n != 4) || prog->engines[0] != WDG_E_COLLIDER || prog->engines[1] != WDG_E_CONTACTLOOP ||
      (n == 4 && prog->engines[2] != WDG_E_INTRAFORCE) || prog->engines[n - 1] != WDG_E_LEAPFROG)
    throw Error(WDG_E_ARG, "graph replay runs the engine list [Collider, ContactLoop, Leapfrog] or [Collider, "
                           "ContactLoop, IntraForce, Leapfrog]");
  if (!prog->leapfrog.reset) throw Error(WDG_E_ARG, "graph replay needs Leapfrog.reset");
}

static void plain_step(Ctx& c, const wdg_program* prog, int64_t step, double dt) {
  wdg_contact_params cp = prog->contact;
  cp.step = step;
  cp.dt = dt;
  contact_loop(c, &cp);
  if (has_intra(prog)) intra_force(c);
  wdg_leapfrog_params lp = prog->leapfrog;
  lp.dt = dt;
  leapfrog(c, &lp);
}

// host state a captured (and so not executed) sequence of launches changes
struct HostState {
  int rcur, lfParity, dirtyIdx;
  bool forcesPending, forceClean, dirtyKnown;
  bool sidxValid[2];
  bool eidxValid, eidxClean;
};
static HostState save(const Ctx& c) {
  return HostState{c.rcur, c.lfParity, c.dirtyIdx, c.forcesPending, c.forceClean, c.dirtyKnown,
                   {c.R[0].sidxValid, c.R[1].sidxValid}, c.eidxValid, c.eidxClean};
}
static void restore(Ctx& c, const HostState& h) {
  c.rcur = h.rcur; c.lfParity = h.lfParity; c.dirtyIdx = h.dirtyIdx;
  c.forcesPending = h.forcesPending; c.forceClean = h.forceClean; c.dirtyKnown = h.dirtyKnown;
  c.R[0].sidxValid = h.sidxValid[0]; c.R[1].sidxValid = h.sidxValid[1];
  c.eidxValid = h.eidxValid;
  c.eidxClean = h.eidxClean;
}

static void capture(Ctx& c, const wdg_program* prog, int K, int64_t step, double dt) {
  auto& G = c.graph;
  if (G.exec) {
    WDG_HIP(hipGraphExecDestroy(G.exec));
    G.exec = nullptr;
  }
  // every buffer a captured launch touches is sized now (no allocation while capturing)
  const int64_t nc = c.con().n;
  c.R[0].ensure(nc);
  c.R[1].ensure(nc);
  if (rolling(prog->contact)) {
    c.R[0].ensureRoll(nc);
    c.R[1].ensureRoll(nc);
  }
  c.act.ensure(ngroups(nc) * GROUP);
  c.desc.ensure(ngroups(nc) * GROUP);
  if (c.field.periodic) c.descDist.ensure(ngroups(nc) * GROUP);
  if (c.trackEnergy)
    c.enPart.ensure(std::max<size_t>((size_t)4 * ngroups(nc) * c.tune.cfBlocks, (size_t)3 * nblocks(c.n, 256)));
  if (energy_per_scene(c)) energy_scenes_prepare(c, nc);
  const HostState h = save(c);
  hipGraph_t graph = nullptr;
  c.graphCapture = true;
  WDG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int j = 0; j < K; j++) {
      // the dirty word the previous step's leapfrog sets
      const int prevWord = DIRTY_PP + (1 - c.lfParity);
      hipLaunchKernelGGL(k_gate, dim3(1), dim3(64), 0, c.stream, c.gstate.p, (const uint32_t*)c.dflags.p, prevWord);
      WDG_LAUNCH_CHECK();
      plain_step(c, prog, step + j, dt);
    }
  } catch (...) {
    c.graphCapture = false;
    (void)hipStreamEndCapture(c.stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    restore(c, h);
    throw;
  }
  c.graphCapture = false;
  WDG_HIP(hipStreamEndCapture(c.stream, &graph));
  WDG_HIP(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0));
  WDG_HIP(hipGraphDestroy(graph));
  restore(c, h);
  G.gen = c.gen;
  G.K = K;
  G.cfBlocks = c.tune.cfBlocks;
  G.track = c.trackEnergy;
  G.rcur = c.rcur;
  G.parity = c.lfParity;
  G.dt = dt;
  G.damping = prog->leapfrog.damping;
  G.cp = prog->contact;
  G.intra = has_intra(prog);
  G.epochs = c.matEpochs;
  G.updAll = c.updatePhys != 0;
  G.evalStress = c.evalStress;
}

static bool graph_valid(const Ctx& c, const wdg_program* prog, int K, double dt) {
  const auto& G = c.graph;
  const wdg_contact_params& a = G.cp;
  const wdg_contact_params& b = prog->contact;
  return G.exec && G.gen == c.gen && G.K == K && G.cfBlocks == c.tune.cfBlocks && G.track == c.trackEnergy &&
         G.intra == has_intra(prog) && G.epochs == c.matEpochs && G.updAll == (c.updatePhys != 0) &&
         G.evalStress == c.evalStress &&
         G.rcur == c.rcur &&
         G.parity == c.lfParity && G.dt == dt && G.damping == prog->leapfrog.damping && a.physKind == b.physKind &&
         a.lawKind == b.lawKind && a.applyForces == b.applyForces && a.trsfRenorm == b.trsfRenorm &&
         a.poisson == b.poisson && (a.en == b.en || (a.en != a.en && b.en != b.en)) && a.noSlip == b.noSlip &&
         a.noBreak == b.noBreak && a.noFrict == b.noFrict && a.noAttraction == b.noAttraction &&
         a.relRollStiff == b.relRollStiff && a.relTwistStiff == b.relTwistStiff && a.rollTanPhi == b.rollTanPhi;
}

// one batch of R replays of the K-step graph (up to R*K steps); returns the steps executed. The
// executed count and the dirty/error words come back in one round trip, and the collider call
// that follows uses them without reading again.
static int replay(Ctx& c, const wdg_program* prog, int K, int R, int64_t step, double dt) {
  // the captured filters keep the per-entry state indices of a clean or current array
  if (!c.eidxValid && !c.eidxClean && c.con().n > 0) {
    eidx_clear(c);
    c.eidxClean = true;
  }
  if (!graph_valid(c, prog, K, dt)) capture(c, prog, K, step, dt);
  uint32_t w[7] = {0u, 0u, 0u, (uint32_t)((uint64_t)step & 0xFFFFFFFFu), (uint32_t)((uint64_t)step >> 32),
                   (uint32_t)std::max(prog->contact.trsfRenorm, 0), 1u};
  WDG_HIP(hipMemcpyAsync(c.gstate.p, w, sizeof(w), hipMemcpyHostToDevice, c.stream));
  for (int r = 0; r < R; r++) WDG_HIP(hipGraphLaunch(c.graph.exec, c.stream));
  WDG_HIP(hipMemcpyAsync(c.hflags + 32, c.gstate.p + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipMemcpyAsync(c.hflags, c.dflags.p, (DIRTY_PP + 2) * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  const int e = (int)c.hflags[32];
  c.hflagsFresh = true;
  // host state after e steps of contact loop + leapfrog
  if (e > 0) {
    if (c.con().n > 0) {
      // the real sets the executed filters wrote carry their per-slot state indices
      c.R[1 - c.rcur].sidxValid = true;
      if (e > 1) c.R[c.rcur].sidxValid = true;
      // the last executed filter wrote the per-entry state indices of the set now current
      c.eidxValid = ngroups(c.con().n) * GROUP < (int64_t)EIDX_PA;
      c.eidxClean = false;
      c.rcur = (e & 1) ? 1 - c.rcur : c.rcur;
    }
    const int lastParity = (c.lfParity + e - 1) & 1;
    c.lfParity = (c.lfParity + e) & 1;
    c.dirtyIdx = DIRTY_PP + lastParity;
    c.dirtyKnown = true;
    c.forcesPending = false;
    c.forceClean = true;
  }
  return e;
}

static void run_graph_steps(Ctx& c, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt) {
  check_program(prog);
  c.gstate.ensure(16);
  int64_t done = 0;
  const int K = c.tune.graphSteps;
  while (done < nSteps) {
    const int64_t nr = c.nRefresh;
    collide(c, &prog->collide);
    if (c.nRefresh != nr) {
      // the run's refresh interval, a running mean over the last few refreshes
      if (c.lastRefreshStep >= 0) {
        const double iv = (double)(*step - c.lastRefreshStep);
        c.refreshEvery = c.refreshEvery > 0 ? 0.75 * c.refreshEvery + 0.25 * iv : iv;
      }
      c.lastRefreshStep = *step;
    }
    // a replay starts from a state its capture saw: forces at their reset value, the
    // contact loop's forces gathered; otherwise one plain step
    const int64_t left = nSteps - done;
    // (a periodic scene's step 0 is a plain one when tracking energy: its leapfrog adds no terms,
    // and a graph captured there would skip them at every replay)
    // (updatePhys 'once' is one plain step, after which the loop reads 'never')
    if (!c.forceClean || c.forcesPending || left < K || c.n == 0 || c.updatePhys == 2 ||
        (c.trackEnergy && c.field.periodic && *step == 0)) {
      plain_step(c, prog, *step, *dt);
      *step += 1;
      done += 1;
      continue;
    }
    // replays per batch: up to the next refresh the run's own refresh interval predicts (a
    // batch that reaches a refresh earlier wastes only the launches of its remaining replays)
    int R = 1;
    if (c.refreshEvery > 0) {
      const double toRefresh = c.refreshEvery - (double)(*step - c.lastRefreshStep);
      R = (int)std::max(1., std::floor(toRefresh / K));
    }
    R = (int)std::min<int64_t>(std::min<int64_t>(R, left / K), 64);
    const int e = replay(c, prog, K, R, *step, *dt);
    if (e == 0) throw Error(WDG_E_STATE, "graph replay executed no step");
    *step += e;
    done += e;
  }
}

void run_graph(Ctx& c, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt) {
  // the host mirror of the flag words serves only the collider call right after a batch
  try {
    run_graph_steps(c, prog, nSteps, step, dt);
  } catch (...) {
    c.hflagsFresh = false;
    throw;
  }
  c.hflagsFresh = false;
}

// ------------------------------------------------------------------ autotuner
int64_t tune(Ctx& c, const wdg_program* prog, int64_t* step, double* dt) {
  check_program(prog);
  const int cfCand[] = {1, 2, 4};
  const int kCand[] = {4, 8, 16};
  wdg_tuning best = c.tune;
  double bestT = INFINITY;
  int64_t taken = 0;
  // warm-up: one plain step so that forces are at their reset value
  run_graph(c, prog, 2, step, dt);
  taken += 2;
  // batched scenes: the per-scene skin first, timed with its refreshes over windows that hold
  // several of them (a larger skin means fewer refreshes but more candidate slots)
  if (c.nScenes > 1) {
    // 16 first: another skin must be faster by 3 % to replace it (at the start, before the
    // scenes move, the skins do not differ and the window's noise would decide)
    const int sCand[] = {16, 8, 32, 0};
    constexpr int W = 96;
    int bestS = c.tune.skinSteps;
    double bestST = INFINITY;
    for (int sk : sCand) {
      c.tune.skinSteps = sk;
      c.boundsValid = false;  // the next collider call refreshes with this skin
      run_graph(c, prog, best.graphSteps, step, dt);
      taken += best.graphSteps;
      WDG_HIP(hipStreamSynchronize(c.stream));
      const auto t0 = std::chrono::steady_clock::now();
      run_graph(c, prog, W, step, dt);
      WDG_HIP(hipStreamSynchronize(c.stream));
      const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / W;
      taken += W;
      if (t < 0.97 * bestST) {
        bestST = t;
        bestS = sk;
      }
    }
    c.tune.skinSteps = bestS;
    best.skinSteps = bestS;
    c.boundsValid = false;
  }
  for (int cf : cfCand)
    for (int K : kCand) {
      c.tune = wdg_tuning{cf, K, best.skinSteps};
      // the first replay captures; the timed ones replay. The refreshes inside the timed
      // window are taken out of its time, so that a window that happens to hold one does not
      // lose to one that does not
      run_graph(c, prog, K, step, dt);
      taken += K;
      const int reps = 3;
      WDG_HIP(hipStreamSynchronize(c.stream));
      const double r0 = c.refreshSec;
      const auto t0 = std::chrono::steady_clock::now();
      run_graph(c, prog, (int64_t)reps * K, step, dt);
      WDG_HIP(hipStreamSynchronize(c.stream));
      const double t = (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() -
                        (c.refreshSec - r0)) /
                       (reps * K);
      taken += (int64_t)reps * K;
      if (t < bestT) {
        bestT = t;
        best = c.tune;
      }
    }
  c.tune = best;
  return taken;
}

}  // namespace wdg
