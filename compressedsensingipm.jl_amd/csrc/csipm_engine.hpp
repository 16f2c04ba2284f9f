// Batch engine of libcsipm: the interior-point driver and its CG over any set
// of operator handles at once.  Host code, part of csipm.hip's translation
// unit (included inside its anonymous namespace, after the pass launchers).
//
// Instances live in slot ranges of one pool of spectrum-ordered vectors (see
// csipm_batch.hpp); every kernel launch covers all instances it is given, so
// a batch of uneven 1D/2D/3D problems costs the launches of one.  The
// operator apply of the batch is a chain of passes per dimensionality: each
// pass is one launch over the tiles of all instances of that dimensionality
// (split only where instances need different workgroup sizes), reading its
// per-instance arguments from a device table.  A single solve is a batch of
// one.  Instances are admitted and retired between interior-point iterations
// (continuous batching): their slots, partial sums and state are reused.

// ---------------------------------------------------------------------------
// beta-layout <-> u conversions of one N-vector of instance h
// sign = 1 for a vector, 0 for a diagonal matrix (whose entries keep their sign)
// ---------------------------------------------------------------------------
template <typename T>
int to_u(csipm_op* h, const T* beta, T* u, hipStream_t st, int sign = 1) {
  const int grid = (int)((h->Nc + kNT - 1) / kNT);
  cpx<T>* uu = (cpx<T>*)u;
  if (h->dim == 1)
    hipLaunchKernelGGL((k_beta_to_u<1, T>), dim3(grid), dim3(kNT), 0, st, beta, uu, h->g, h->H1, h->Nc, sign);
  else if (h->dim == 2)
    hipLaunchKernelGGL((k_beta_to_u<2, T>), dim3(grid), dim3(kNT), 0, st, beta, uu, h->g, h->H1, h->Nc, sign);
  else
    hipLaunchKernelGGL((k_beta_to_u<3, T>), dim3(grid), dim3(kNT), 0, st, beta, uu, h->g, h->H1, h->Nc, sign);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

template <typename T>
int from_u(csipm_op* h, const T* u, T* beta, hipStream_t st, int sign = 1) {
  const int grid = (int)((h->Nc + kNT - 1) / kNT);
  const cpx<T>* uu = (const cpx<T>*)u;
  if (h->dim == 1)
    hipLaunchKernelGGL((k_u_to_beta<1, T>), dim3(grid), dim3(kNT), 0, st, uu, beta, h->g, h->H1, h->Nc, sign);
  else if (h->dim == 2)
    hipLaunchKernelGGL((k_u_to_beta<2, T>), dim3(grid), dim3(kNT), 0, st, uu, beta, h->g, h->H1, h->Nc, sign);
  else
    hipLaunchKernelGGL((k_u_to_beta<3, T>), dim3(grid), dim3(kNT), 0, st, uu, beta, h->g, h->H1, h->Nc, sign);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// operator chains
// ---------------------------------------------------------------------------
// MTM: the fused M^T M pass sequence on u (with the epilogue in the last
// pass); MTZ: M^T z from real samples into u
enum { CH_MTM = 0, CH_MTZ = 1 };

// instances whose chains have the same pass sequence share launches
int chain_key(const csipm_op* h) { return h->big ? (h->obs ? 11 : 10) : h->dim; }

int chain_specs(const csipm_op* h, int which, PassSpec* s) {
  if (h->big) return big_specs(h, which == CH_MTM ? BIG_MTM : BIG_R2C, SRC_U, DST_U, s);
  const int dim = h->dim;
  if (which == CH_MTM) {
    if (dim == 1) { s[0] = {PK_DIM1, 1, SRC_U, DST_U, MODE_MTM, 0}; return 1; }
    if (dim == 2) {
      s[0] = {PK_LINES, 2, SRC_U, DST_BUF, 0, 1};
      s[1] = {PK_DIM1, 2, SRC_BUF, DST_BUF, MODE_MTM, 0};
      s[2] = {PK_LINES, 2, SRC_BUF, DST_U, 1, 1};
      return 3;
    }
    s[0] = {PK_LINES, 3, SRC_U, DST_BUF, 0, 2};
    s[1] = {PK_LINES, 3, SRC_BUF, DST_BUF, 0, 1};
    s[2] = {PK_DIM1, 3, SRC_BUF, DST_BUF, MODE_MTM, 0};
    s[3] = {PK_LINES, 3, SRC_BUF, DST_BUF, 1, 1};
    s[4] = {PK_LINES, 3, SRC_BUF, DST_U, 1, 2};
    return 5;
  }
  if (dim == 1) { s[0] = {PK_DIM1, 1, SRC_REAL, DST_U, MODE_R2C, 0}; return 1; }
  if (dim == 2) {
    s[0] = {PK_DIM1, 2, SRC_REAL, DST_BUF, MODE_R2C, 0};
    s[1] = {PK_LINES, 2, SRC_BUF, DST_U, 1, 1};
    return 2;
  }
  s[0] = {PK_DIM1, 3, SRC_REAL, DST_BUF, MODE_R2C, 0};
  s[1] = {PK_LINES, 3, SRC_BUF, DST_BUF, 1, 1};
  s[2] = {PK_LINES, 3, SRC_BUF, DST_U, 1, 2};
  return 3;
}

// the batched kernel of one pass spec (the combinations the chains use)
template <typename T, int NT>
int launch_spec(const PassSpec& s, int grid, size_t lds, hipStream_t st, const void* tabv,
                BSeg seg) {
  const PassArgs<T>* tab = (const PassArgs<T>*)tabv;
#define CSIPM_LAUNCH(kern)                                                      \
  do {                                                                          \
    CHK(allow_lds(kern, lds));                                                  \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, st, tab, seg);          \
    HIPCHK(hipGetLastError());                                                  \
    return CSIPM_OK;                                                            \
  } while (0)
  if (s.kind == PK_LINES) {
    if constexpr (NT == kNT || NT == kNTSmall) {
      if (s.D == 2 && s.src == SRC_U && s.dst == DST_BUF && !s.fm)
        CSIPM_LAUNCH((kb_lines<T, NT, kE, 2, SRC_U, DST_BUF, false>));
      if (s.D == 2 && s.src == SRC_BUF && s.dst == DST_U && s.fm)
        CSIPM_LAUNCH((kb_lines<T, NT, kE, 2, SRC_BUF, DST_U, true>));
      if (s.D == 3 && s.src == SRC_U && s.dst == DST_BUF && !s.fm)
        CSIPM_LAUNCH((kb_lines<T, NT, kE, 3, SRC_U, DST_BUF, false>));
      if (s.D == 3 && s.src == SRC_BUF && s.dst == DST_BUF && !s.fm)
        CSIPM_LAUNCH((kb_lines<T, NT, kE, 3, SRC_BUF, DST_BUF, false>));
      if (s.D == 3 && s.src == SRC_BUF && s.dst == DST_BUF && s.fm)
        CSIPM_LAUNCH((kb_lines<T, NT, kE, 3, SRC_BUF, DST_BUF, true>));
      if (s.D == 3 && s.src == SRC_BUF && s.dst == DST_U && s.fm)
        CSIPM_LAUNCH((kb_lines<T, NT, kE, 3, SRC_BUF, DST_U, true>));
    }
    return fail(CSIPM_ERR_UNSUPPORTED, "line pass with %d threads", NT);
  } else if (s.kind == PK_COL) {
    if constexpr (NT != kNT) return fail(CSIPM_ERR_UNSUPPORTED, "column pass with %d threads", NT);
    if (s.fm == C_FIRST) CSIPM_LAUNCH((kb_col<T, kNT, kE, SRC_U, DST_BUF, C_FIRST>));
    if (s.fm == C_REAL_IN) CSIPM_LAUNCH((kb_col<T, kNT, kE, SRC_REAL, DST_BUF, C_REAL_IN>));
    if (s.fm == C_MID) CSIPM_LAUNCH((kb_col<T, kNT, kE, SRC_BUF, DST_BUF, C_MID>));
    if (s.fm == C_LAST) CSIPM_LAUNCH((kb_col<T, kNT, kE, SRC_BUF, DST_U, C_LAST>));
  } else if (s.kind == PK_ROW) {
    if constexpr (NT != kNT) return fail(CSIPM_ERR_UNSUPPORTED, "row pass with %d threads", NT);
    if (s.fm == R_MTM) CSIPM_LAUNCH((kb_row<T, kNT, kE, R_MTM>));
    if (s.fm == R_R2C) CSIPM_LAUNCH((kb_row<T, kNT, kE, R_R2C>));
    if (s.fm == R_CONV_INV) CSIPM_LAUNCH((kb_row<T, kNT, kE, R_CONV_INV>));
    if (s.fm == R_CONV_FWD) CSIPM_LAUNCH((kb_row<T, kNT, kE, R_CONV_FWD>));
  } else {
    if (s.D == 1 && s.src == SRC_U && s.fm == MODE_MTM)
      CSIPM_LAUNCH((kb_dim1<T, NT, kE, 1, SRC_U, DST_U, MODE_MTM>));
    if (s.D == 1 && s.src == SRC_REAL && s.fm == MODE_R2C)
      CSIPM_LAUNCH((kb_dim1<T, NT, kE, 1, SRC_REAL, DST_U, MODE_R2C>));
    if (s.D == 2 && s.src == SRC_BUF && s.fm == MODE_MTM)
      CSIPM_LAUNCH((kb_dim1<T, NT, kE, 2, SRC_BUF, DST_BUF, MODE_MTM>));
    if (s.D == 2 && s.src == SRC_REAL && s.fm == MODE_R2C)
      CSIPM_LAUNCH((kb_dim1<T, NT, kE, 2, SRC_REAL, DST_BUF, MODE_R2C>));
    if (s.D == 3 && s.src == SRC_BUF && s.fm == MODE_MTM)
      CSIPM_LAUNCH((kb_dim1<T, NT, kE, 3, SRC_BUF, DST_BUF, MODE_MTM>));
    if (s.D == 3 && s.src == SRC_REAL && s.fm == MODE_R2C)
      CSIPM_LAUNCH((kb_dim1<T, NT, kE, 3, SRC_REAL, DST_BUF, MODE_R2C>));
  }
#undef CSIPM_LAUNCH
  return fail(CSIPM_ERR_UNSUPPORTED, "no batched kernel for pass (%d,%d,%d,%d,%d)", s.kind, s.D,
              s.src, s.dst, s.fm);
}

// Launch tables of a rebuild (chain pass arguments, block -> instance maps,
// vector segments) are staged on the host, packed into one pinned buffer and
// copied to one device buffer by a single asynchronous copy on the engine's
// stream.  Both buffers are kept and reused by the next rebuild: the copy is
// ordered after every earlier launch that reads the old tables, and the host
// buffer is reused once the event after the previous copy has passed, so a
// rebuild allocates, frees and synchronises nothing unless the tables
// outgrow the buffers.
struct TableArena {
  char* dev = nullptr;
  char* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;      // recorded after the last copy out of `host`
  std::vector<char> stage;      // the tables of the rebuild in progress
  void begin() { stage.clear(); }
  // appends bytes (256-byte aligned); returns their offset
  size_t put(const void* p, size_t bytes) {
    const size_t off = (stage.size() + 255) & ~(size_t)255;
    stage.resize(off + bytes);
    if (bytes) std::memcpy(stage.data() + off, p, bytes);
    return off;
  }
  // copies the staged tables to the device; *base: their device address
  int commit(hipStream_t st, char** base) {
    const size_t need = std::max<size_t>(stage.size(), 256);
    if (ev) HIPCHK(hipEventSynchronize(ev));
    if (need > cap) {
      // launches still reading the old tables finish before the buffer goes
      HIPCHK(hipStreamSynchronize(st));
      dfree(dev);
      hfree(host);
      dev = nullptr;
      host = nullptr;
      cap = 0;
      const size_t want = (size_t)1 << block_class(std::max(need, (size_t)1 << 14) * 2);
      HIPCHK(dmalloc((void**)&dev, want));
      HIPCHK(hmalloc((void**)&host, want));
      cap = want;
    }
    if (!ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    std::memcpy(host, stage.data(), stage.size());
    HIPCHK(hipMemcpyAsync(dev, host, stage.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(ev, st));
    *base = dev;
    return CSIPM_OK;
  }
  ~TableArena() {
    if (ev) {
      (void)hipEventSynchronize(ev);
      (void)hipEventDestroy(ev);
    }
    dfree(dev);
    hfree(host);
  }
};

// one launch of a chain: a pass over a subset of instances
struct ChainSub {
  PassSpec spec;
  int pass = 0;          // index of the pass in its chain
  std::vector<int> ids;  // member instance ids
  int nt = kNT, grid = 0;
  size_t lds = 0;
  void* tab = nullptr;   // device PassArgs<T>[members] (in a table arena)
  int* seg = nullptr;    // device: block -> member, then member -> first block
  size_t tab_off = 0, seg_off = 0;   // their offsets in the arena
  int members = 0;
};

struct Chain {
  int esize = 8;
  std::vector<ChainSub> subs;
  void release() { subs.clear(); }   // the tables live in a table arena
  // device addresses of the tables once the arena is committed
  void bind(char* base) {
    for (ChainSub& s : subs) {
      s.tab = base + s.tab_off;
      s.seg = (int*)(base + s.seg_off);
    }
  }
};

// a vector segment staged in a table arena: block -> instance, instance ->
// first block (maxinst + 1 entries), then the id list
struct VsegOff {
  size_t off = 0;
  int nblk = 0, nfirst = 0, grid = 0;
  void bind(char* base, BSeg* seg, const int** ids) const {
    int* m = (int*)(base + off);
    seg->inst = m;
    seg->first = m + nblk;
    *ids = m + nblk + nfirst;
  }
};

template <typename T>
int launch_sub(const ChainSub& s, hipStream_t st) {
  BSeg seg{s.seg, s.seg + s.grid};
  if (s.nt == kNT) return launch_spec<T, kNT>(s.spec, s.grid, s.lds, st, s.tab, seg);
  if (s.nt == kNTSmall) return launch_spec<T, kNTSmall>(s.spec, s.grid, s.lds, st, s.tab, seg);
  return launch_spec<T, kNTBig>(s.spec, s.grid, s.lds, st, s.tab, seg);
}

template <typename T>
int run_chain(const Chain& c, hipStream_t st) {
  for (const ChainSub& s : c.subs) CHK(launch_sub<T>(s, st));
  return CSIPM_OK;
}

// the same with the passes of the 1D and 2D instances on `side`, concurrent
// with those of the 3D instances on `st` (the instances of a launch are
// independent; each dimensionality keeps its pass order): a fork from `st`
// and a join back into it, which a graph capture of `st` records as two
// branches.  The small instances' passes, latency-bound launches of few
// workgroups, then fill the CUs the 3D passes leave idle.
template <typename T>
int run_chain_par(const Chain& c, hipStream_t st, hipStream_t side, hipEvent_t fork_ev,
                  hipEvent_t join_ev) {
  bool big = false, small = false;
  for (const ChainSub& s : c.subs) (s.spec.D == 3 ? big : small) = true;
  if (!big || !small || !side) return run_chain<T>(c, st);
  HIPCHK(hipEventRecord(fork_ev, st));
  HIPCHK(hipStreamWaitEvent(side, fork_ev, 0));
  for (const ChainSub& s : c.subs)
    if (s.spec.D != 3) CHK(launch_sub<T>(s, side));
  HIPCHK(hipEventRecord(join_ev, side));
  for (const ChainSub& s : c.subs)
    if (s.spec.D == 3) CHK(launch_sub<T>(s, st));
  HIPCHK(hipStreamWaitEvent(st, join_ev, 0));
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// the engine
// ---------------------------------------------------------------------------
constexpr int64_t kSlotAlign = 256;   // slot ranges start on 256-slot boundaries
constexpr int kPartPerAlign = 64;     // partial-sum slots per 256 vector slots (a line pass
                                      // over U slots has at most U/4 workgroups)

// FP64 control vectors (one block, pool length) and two-block FP64 vectors
enum { V_B = 0, V_BETA, V_C, V_Z1, V_Z2, V_QB, V_S1, V_S2, V_DT1, V_DT2, V_DZ1, V_DZ2, V_DGD,
       V_NV1 };
enum { V2_RHSD = 0, V2_DD, V2_TMP, V_NV2 };
// CG-precision vectors: one block (L1 or D, L2, P11 or P, P12, P22), two-block
// (rhs, d) and the CG vectors r, z, p, q (two-block)
enum { C_L1 = 0, C_L2, C_P11, C_P12, C_P22, C_NV1 };
enum { C2_RHS = 0, C2_D, C2_R, C2_Z, C2_P, C2_Q, C_NV2 };

// chains of the interior-point iteration
enum { CALL_QB = 0, CALL_REF, CALL_CGA, CALL_CGR, CALL_MTZ, CALL_N };

struct Engine {
  int tc_size = 8;          // CG / operator precision of the CG
  bool mixed = false;       // reduced system, FP32 CG, FP64 refinement
  hipStream_t stream = nullptr;   // the engine's own stream (all its work)
  hipStream_t user = nullptr;     // the caller's stream, ordered around every call
  hipEvent_t ev = nullptr;
  // a second stream: an apply's passes of the 1D and 2D instances run on it
  // beside those of the 3D instances (forked from and joined back into
  // `stream`, also inside the captured CG chunks)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool use_graphs = true;         // CG chunks replayed as captured HIP graphs
  // captured CG chunks: [0] four plain iterations, [1] three plain + one with
  // the reliable update; valid for (graph_nb, graph_rel, graph_version)
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_nb = 0, graph_rel = -1;
  int64_t graph_version = -1;
  int64_t tables_version = 0;
  int cg_on_tables = 0;            // CG solves run on the current tables
  std::vector<int64_t> table_sig;  // (id, admission, offset) of the set the tables were built for
  int64_t admissions = 0;
  int table_nb = 0;
  bool table_ipm = false;
  // host-side seconds and counts, for csipm_batch_stats
  enum { ST_STEP = 0, ST_TABLES, ST_CAPTURE, ST_INIT, ST_CG, ST_NTABLES, ST_NCAPTURE, ST_NSTEP,
         ST_N };
  double stats[ST_N] = {};
  int64_t cap = 0;          // pool slots
  int maxinst = 0;
  struct Inst {
    csipm_op* h = nullptr;
    int64_t off = 0, len = 0;
    bool live = false;      // admitted and not yet retired
    bool fresh = false;     // admitted, not yet initialised
    double lam = 0;
    const double* z0 = nullptr;
    double* beta_out = nullptr;
    double* c_out = nullptr;  // optional: the final c iterate (reference order)
    int64_t serial = 0;       // admission number (unique over the engine's life)
  };
  std::vector<Inst> inst;
  std::vector<std::pair<int64_t, int64_t>> freelist;   // (offset, length)
  // device state
  std::vector<BInst> info_h;   // host mirror of info
  BInst* info = nullptr;
  BFlags* flags = nullptr;
  BFlags* flags_h = nullptr;
  CgState* cg = nullptr;
  CgState* cg_h = nullptr;
  double* scal = nullptr;
  double* part = nullptr;
  uint32_t* valid = nullptr;
  double* Wd = nullptr;      // V_NV1 * cap + V_NV2 * 2 cap doubles
  void* Wc = nullptr;        // C_NV1 * cap + C_NV2 * 2 cap CG-precision elements
  // launch tables of the current active set
  std::vector<int> active;   // instance ids
  TableArena tabs;           // the active set's tables
  TableArena init_tabs;      // the tables of the fresh-instance initialisation
  TableArena rec_tabs;       // staged instance records (info_h) and new ids
  std::vector<int> pending;  // admitted since the last record upload (flags to zero)
  bool records_dirty = false;
  bool have_tables = false;
  int vgrid = 0;
  BSeg vseg{};
  const int* ids = nullptr;
  Chain chains[CALL_N];
  bool tables_dirty = true;

  double* W(int k) { return Wd + (size_t)k * cap; }
  double* W2(int k) { return Wd + (size_t)V_NV1 * cap + (size_t)k * 2 * cap; }
  template <typename T>
  T* C(int k) { return (T*)Wc + (size_t)k * cap; }
  template <typename T>
  T* C2(int k) { return (T*)Wc + (size_t)C_NV1 * cap + (size_t)k * 2 * cap; }

  void release_tables() {
    for (Chain& c : chains) c.release();
    have_tables = false;
  }
  void release_graphs() {
    for (auto& g : graph)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    graph_version = -1;
  }
  ~Engine() {
    for (Inst& in : inst)
      if (in.h) in.h->admitted.fetch_sub(1);
    // blocks (tables, staging, vectors) go back to the caches, and graphs are
    // destroyed: the streams' work ends first
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    release_graphs();
    release_tables();
    if (stream) (void)hipStreamDestroy(stream);
    if (side) (void)hipStreamDestroy(side);
    if (ev) (void)hipEventDestroy(ev);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    for (void* p : {(void*)info, (void*)flags, (void*)cg, (void*)scal, (void*)part, (void*)valid,
                    (void*)Wd, Wc})
      dfree(p);
    hfree(flags_h);
    hfree(cg_h);
  }
};

struct StatTimer {
  double* slot;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit StatTimer(double* s) : slot(s) {}
  ~StatTimer() {
    *slot += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

int64_t slots_of(const csipm_op* h) { return (h->U + kSlotAlign - 1) / kSlotAlign * kSlotAlign; }

// device bytes of the engine pool per slot
size_t engine_bytes_per_slot(int tc_size) {
  return sizeof(double) * (V_NV1 + 2 * V_NV2) + (size_t)tc_size * (C_NV1 + 2 * C_NV2) +
         sizeof(double) * kPartPerAlign / kSlotAlign + 1;
}

int engine_create(Engine** out, int64_t cap_slots, int maxinst, int tc_size, bool mixed,
                  hipStream_t st) {
  Engine* E = new (std::nothrow) Engine();
  if (!E) return fail(CSIPM_ERR_ALLOC, "out of host memory");
  E->tc_size = tc_size;
  E->mixed = mixed;
  E->user = st;
  E->use_graphs = std::getenv("CSIPM_NO_GRAPH") == nullptr;
  E->cap = (cap_slots + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
  E->maxinst = maxinst;
  E->inst.resize((size_t)maxinst);
  E->info_h.assign((size_t)maxinst, BInst{});
  E->freelist.push_back({0, E->cap});
  auto bad = [&](int rc) { delete E; return rc; };
#define ECHK(expr)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return bad(fail(CSIPM_ERR_ALLOC, "%s failed: %s", #expr, hipGetErrorString(e_))); \
  } while (0)
  ECHK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
  ECHK(hipStreamCreateWithFlags(&E->side, hipStreamNonBlocking));
  ECHK(hipEventCreateWithFlags(&E->ev, hipEventDisableTiming));
  ECHK(hipEventCreateWithFlags(&E->fork_ev, hipEventDisableTiming));
  ECHK(hipEventCreateWithFlags(&E->join_ev, hipEventDisableTiming));
  ECHK(dmalloc((void**)&E->info, sizeof(BInst) * (size_t)maxinst));
  ECHK(dmalloc((void**)&E->flags, sizeof(BFlags) * (size_t)maxinst));
  ECHK(hmalloc((void**)&E->flags_h, sizeof(BFlags) * (size_t)maxinst));
  ECHK(dmalloc((void**)&E->cg, sizeof(CgState) * (size_t)maxinst));
  ECHK(hmalloc((void**)&E->cg_h, sizeof(CgState) * (size_t)maxinst));
  ECHK(dmalloc((void**)&E->scal, sizeof(double) * S_NSLOTS * (size_t)maxinst));
  ECHK(dmalloc((void**)&E->part, sizeof(double) * (size_t)(E->cap / kSlotAlign * kPartPerAlign)));
  ECHK(dmalloc((void**)&E->valid, sizeof(uint32_t) * (size_t)(E->cap / 32)));
  ECHK(dmalloc((void**)&E->Wd, sizeof(double) * (size_t)E->cap * (V_NV1 + 2 * V_NV2)));
  ECHK(dmalloc(&E->Wc, (size_t)tc_size * (size_t)E->cap * (C_NV1 + 2 * C_NV2)));
#undef ECHK
  *out = E;
  return CSIPM_OK;
}

// order the engine's stream after the caller's work, and the caller's after ours
int engine_enter(Engine* E) {
  CHK(no_capture(E->user, "a batch-engine call"));
  HIPCHK(hipEventRecord(E->ev, E->user));
  HIPCHK(hipStreamWaitEvent(E->stream, E->ev, 0));
  return CSIPM_OK;
}
int engine_leave(Engine* E) {
  HIPCHK(hipEventRecord(E->ev, E->stream));
  HIPCHK(hipStreamWaitEvent(E->user, E->ev, 0));
  return CSIPM_OK;
}

// first-fit slot allocation
bool slots_alloc(Engine* E, int64_t len, int64_t* off) {
  for (size_t i = 0; i < E->freelist.size(); ++i) {
    auto& f = E->freelist[i];
    if (f.second >= len) {
      *off = f.first;
      f.first += len;
      f.second -= len;
      if (f.second == 0) E->freelist.erase(E->freelist.begin() + (long)i);
      return true;
    }
  }
  return false;
}

void slots_free(Engine* E, int64_t off, int64_t len) {
  E->freelist.push_back({off, len});
  std::sort(E->freelist.begin(), E->freelist.end());
  std::vector<std::pair<int64_t, int64_t>> merged;
  for (auto& f : E->freelist) {
    if (!merged.empty() && merged.back().first + merged.back().second == f.first)
      merged.back().second += f.second;
    else
      merged.push_back(f);
  }
  E->freelist.swap(merged);
}

// vector-kernel blocks of an instance: 1024 slots per block, at most 2048
int vblocks(int64_t n) {
  int64_t b = (n + 1023) / 1024;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 2048));
}

// launch shapes of instance h inside a chain over `members` instances: the
// workgroup-count target of a single launch is shared among them
struct BatchCfg {
  LineCfg lines[3];
  Dim1Cfg d1;
  LineCfg col;   // four-step passes of long 1D lines
  Dim1Cfg row;
};

template <typename T>
BatchCfg batch_cfg(csipm_op* h, int members) {
  const int target = std::max(1, 1024 / std::max(1, members));
  const size_t cs = sizeof(cpx<T>);
  BatchCfg b{};
  if (h->big) {
    b.col = col_cfg(h->bA, h->bB, cs, target);
    b.row = row_cfg(h->bA, h->bB, cs, target);
    b.d1.nt = kNT;
    return b;
  }
  const TuneVal t0 = tune_lookup(1, cs, h->plan[0].L, h->rows[0]);
  b.d1 = dim1_cfg(h->plan[0].L, h->rows[0], (int)(h->N / h->n[0]), cs, target, t0.lg_cap);
  TuneVal ta[3];
  for (int a = 1; a < h->dim; ++a) {
    ta[a] = tune_lookup(0, cs, h->plan[a].L, h->rows[a]);
    b.lines[a] = line_cfg(h->H1, h->plan[a].L, h->rows[a], h->n[a == 1 ? 2 : 1], cs, target,
                          ta[a].lg_cap);
  }
  // a tile of at most kNTSmall*kE elements runs on kNTSmall threads: a stage
  // of radix R has L*lines/R butterflies, and on 256 threads a 255-point
  // tile of 8 lines (120 radix-17 butterflies) would leave half the lanes
  // idle (a tuned choice replaces the rule where the tile fits it)
  auto small = [](int64_t elems) { return elems <= (int64_t)kNTSmall * kE; };
  auto pick = [&](int nt_rule, int64_t elems, const TuneVal& t) {
    if (nt_rule == kNTBig) return nt_rule;
    if (t.nt == kNTSmall && small(elems)) return kNTSmall;
    if (t.nt == kNT) return kNT;
    return small(elems) ? kNTSmall : kNT;
  };
  b.d1.nt = pick(b.d1.nt, (int64_t)h->plan[0].L << b.d1.lg, t0);
  for (int a = 1; a < h->dim; ++a)
    b.lines[a].nt = pick(kNT, (int64_t)h->plan[a].L << b.lines[a].lg, ta[a]);
  return b;
}

// per-instance epilogue / source of a chain call
template <typename T>
int chain_call_args(Engine* E, int id, int call, int nb, const PassSpec& s,
                    const BatchCfg& bc, PassArgs<T>& a) {
  Engine::Inst& in = E->inst[(size_t)id];
  csipm_op* h = in.h;
  const int64_t off = in.off, n = h->U;
  Epi<T> e{};
  const void* src = nullptr;
  const T* real_in = nullptr;
  const int* done = &E->flags[id].done;
  T scale = (T)(1.0 / (double)h->N);
  const int64_t o2 = 2 * off;
  switch (call) {
    case CALL_QB:
      src = E->W(V_BETA) + off;
      e.kind = 0;
      e.out = (T*)(E->W(V_QB) + off);
      break;
    case CALL_REF:
      src = E->W2(V2_DD) + o2;
      e.kind = 2;
      e.xb = (const T*)(E->W2(V2_DD) + o2);
      e.L1 = (const T*)(E->W(V_DGD) + off);
      e.yb = (T*)(E->W2(V2_TMP) + o2);
      e.out = (T*)(E->W2(V2_TMP) + o2 + n);   // M^T M of the direction, for Q beta
      e.alpha = (T)-1;
      e.beta = (T)1;
      break;
    case CALL_CGA:
    case CALL_CGR: {
      T* x = E->C2<T>(call == CALL_CGA ? C2_P : C2_D) + o2;
      T* q = E->C2<T>(C2_Q) + o2;
      src = x;
      e.kind = nb == 2 ? 1 : 2;
      e.xb = x;
      e.xz = x + n;
      e.L1 = E->C<T>(C_L1) + off;
      e.L2 = E->C<T>(C_L2) + off;
      e.yb = q;
      e.yz = q + n;
      e.alpha = (T)1;
      e.beta = (T)0;
      if (call == CALL_CGA) e.dot_part = E->part + off / kSlotAlign * kPartPerAlign;
      done = &E->cg[id].done;
      break;
    }
    case CALL_MTZ:
      real_in = (const T*)in.z0;
      scale = (T)(1.0 / std::sqrt((double)h->N));
      e.kind = 0;
      e.out = (T*)(E->W(V_B) + off);
      break;
    default:
      return fail(CSIPM_ERR_ARG, "bad chain call");
  }
  if (s.kind == PK_LINES) a = lines_args<T>(h, s.axis, src, e, done, &bc.lines[s.axis]);
  else if (s.kind == PK_DIM1) a = dim1_args<T>(h, src, real_in, nullptr, e, scale, done, &bc.d1);
  else a = big_args<T>(h, s.kind == PK_COL, src, real_in, nullptr, e, scale, done, &bc.col, &bc.row);
  return CSIPM_OK;
}

// workgroup size of a pass of an instance with launch shapes bc
inline int pass_nt(const PassSpec& s, const BatchCfg& bc) {
  if (s.kind == PK_DIM1) return bc.d1.nt;
  if (s.kind == PK_LINES) return bc.lines[s.axis].nt;
  return kNT;
}

template <typename T>
int build_chain(Engine* E, TableArena& A, const std::vector<int>& ids, int call, int nb,
                Chain& out) {
  out.release();
  out.esize = sizeof(T);
  for (int key : {1, 2, 3, 10, 11}) {
    std::vector<int> mem;
    for (int id : ids)
      if (chain_key(E->inst[(size_t)id].h) == key) mem.push_back(id);
    if (mem.empty()) continue;
    PassSpec specs[8];
    const int np = chain_specs(E->inst[(size_t)mem[0]].h, call == CALL_MTZ ? CH_MTZ : CH_MTM,
                               specs);
    for (int p = 0; p < np; ++p) {
      const PassSpec& s = specs[p];
      for (int nt : {kNTSmall, kNT, kNTBig}) {
        std::vector<PassArgs<T>> tab;
        std::vector<int> grids;
        size_t lds = 0;
        for (int id : mem) {
          csipm_op* h = E->inst[(size_t)id].h;
          const BatchCfg bc = batch_cfg<T>(h, (int)ids.size());
          const int want = pass_nt(s, bc);
          if (want != nt) continue;
          PassArgs<T> a{};
          CHK(chain_call_args<T>(E, id, call, nb, s, bc, a));
          tab.push_back(a);
          if (s.kind == PK_LINES) {
            grids.push_back(bc.lines[s.axis].grid);
            lds = std::max(lds, bc.lines[s.axis].lds);
          } else if (s.kind == PK_DIM1) {
            grids.push_back(bc.d1.grid);
            lds = std::max(lds, bc.d1.lds);
          } else if (s.kind == PK_COL) {
            grids.push_back(bc.col.grid);
            lds = std::max(lds, bc.col.lds);
          } else {
            grids.push_back(bc.row.grid);
            lds = std::max(lds, bc.row.lds);
          }
        }
        if (tab.empty()) continue;
        ChainSub sub;
        sub.spec = s;
        sub.pass = p;
        for (int id : mem) {
          csipm_op* hm = E->inst[(size_t)id].h;
          const int want = pass_nt(s, batch_cfg<T>(hm, (int)ids.size()));
          if (want == nt) sub.ids.push_back(id);
        }
        sub.nt = nt;
        sub.lds = lds;
        sub.members = (int)tab.size();
        std::vector<int> segv;
        std::vector<int> first;
        int tot = 0;
        for (size_t m = 0; m < tab.size(); ++m) {
          first.push_back(tot);
          for (int b = 0; b < grids[m]; ++b) segv.push_back((int)m);
          tot += grids[m];
        }
        first.push_back(tot);
        sub.grid = tot;
        segv.insert(segv.end(), first.begin(), first.end());
        sub.tab_off = A.put(tab.data(), sizeof(PassArgs<T>) * tab.size());
        sub.seg_off = A.put(segv.data(), sizeof(int) * segv.size());
        out.subs.push_back(sub);
      }
    }
  }
  return CSIPM_OK;
}

// final-pass partial count of the CG apply of instance h (precision T) in a
// chain over `members` instances
template <typename T>
int cg_final_grid(csipm_op* h, int members) {
  const BatchCfg bc = batch_cfg<T>(h, members);
  if (h->big) return bc.col.grid;
  if (h->dim == 1) return bc.d1.grid;
  return bc.lines[h->dim - 1].grid;
}

// the instance records (info_h) to the device and zeroed flags for the
// instances admitted since the last upload, staged through a table arena:
// asynchronous on the engine's stream, no synchronisation
int engine_upload_records(Engine* E) {
  if (!E->records_dirty && E->pending.empty()) return CSIPM_OK;
  TableArena& A = E->rec_tabs;
  A.begin();
  const size_t ro = A.put(E->info_h.data(), sizeof(BInst) * E->info_h.size());
  const size_t io = A.put(E->pending.data(), sizeof(int) * E->pending.size());
  char* base = nullptr;
  CHK(A.commit(E->stream, &base));
  HIPCHK(hipMemcpyAsync(E->info, base + ro, sizeof(BInst) * E->info_h.size(),
                        hipMemcpyDeviceToDevice, E->stream));
  const int n = (int)E->pending.size();
  if (n > 0) {
    hipLaunchKernelGGL(bk_reset_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, E->stream,
                       (const int*)(base + io), n, E->flags);
    HIPCHK(hipGetLastError());
  }
  E->pending.clear();
  E->records_dirty = false;
  return CSIPM_OK;
}

// BInst.gcg of the active instances for a CG chain over all of them
template <typename TC>
int update_gcg(Engine* E) {
  const int m = (int)E->active.size();
  for (int id : E->active) {
    Engine::Inst& in = E->inst[(size_t)id];
    const int g = cg_final_grid<TC>(in.h, m);
    if (g + (int64_t)vblocks(in.h->U) > in.len / kSlotAlign * kPartPerAlign)
      return fail(CSIPM_ERR_UNSUPPORTED, "partial sums of instance exceed its share");
    E->info_h[(size_t)id].gcg = g;
  }
  E->records_dirty = true;
  return engine_upload_records(E);
}

// vector segment + id list of `ids`, staged in A
void stage_vseg(Engine* E, TableArena& A, const std::vector<int>& ids, VsegOff* out) {
  std::vector<int> blk, first(E->maxinst + 1, 0);
  int tot = 0;
  // first[] is indexed by instance id: ids not in the set get empty ranges
  std::vector<int> nb(E->maxinst, 0);
  for (int id : ids) nb[(size_t)id] = vblocks(E->inst[(size_t)id].h->U);
  for (int id = 0; id < E->maxinst; ++id) {
    first[(size_t)id] = tot;
    for (int b = 0; b < nb[(size_t)id]; ++b) blk.push_back(id);
    tot += nb[(size_t)id];
  }
  first[(size_t)E->maxinst] = tot;
  std::vector<int> all(blk);
  all.insert(all.end(), first.begin(), first.end());
  all.insert(all.end(), ids.begin(), ids.end());
  all.push_back(0);   // never empty
  out->off = A.put(all.data(), sizeof(int) * all.size());
  out->nblk = (int)blk.size();
  out->nfirst = (int)first.size();
  out->grid = tot;
}

// launch tables of the active set: the vector segment and the chains of the
// interior-point iteration (ipm = false: only the CG apply, for a plain CG)
template <typename TC>
int engine_tables(Engine* E, int nb, bool ipm) {
  CHK(engine_upload_records(E));
  std::vector<int64_t> sig;
  for (int id : E->active) {
    // the admission number, not the handle address: a handle created after
    // another was destroyed can reuse its address while its device buffers
    // (scratch, twiddles, mask) live elsewhere
    sig.push_back(id);
    sig.push_back(E->inst[(size_t)id].serial);
    sig.push_back(E->inst[(size_t)id].off);
  }
  if (E->tables_dirty && sig == E->table_sig && nb == E->table_nb && ipm == E->table_ipm &&
      E->have_tables) {
    // the same instances in the same slots: the tables stand (the instance
    // records were rewritten by the admissions)
    CHK(update_gcg<TC>(E));
    E->tables_dirty = false;
  }
  if (!E->tables_dirty) return CSIPM_OK;
  StatTimer tm(&E->stats[Engine::ST_TABLES]);
  E->stats[Engine::ST_NTABLES] += 1;
  E->release_tables();
  E->table_sig.clear();
  TableArena& A = E->tabs;
  A.begin();
  VsegOff vs;
  stage_vseg(E, A, E->active, &vs);
  CHK(build_chain<TC>(E, A, E->active, CALL_CGA, nb, E->chains[CALL_CGA]));
  CHK(build_chain<TC>(E, A, E->active, CALL_CGR, nb, E->chains[CALL_CGR]));
  if (ipm) {
    if (E->mixed) CHK(build_chain<double>(E, A, E->active, CALL_REF, nb, E->chains[CALL_REF]));
    else CHK(build_chain<double>(E, A, E->active, CALL_QB, nb, E->chains[CALL_QB]));
  }
  char* base = nullptr;
  CHK(A.commit(E->stream, &base));
  vs.bind(base, &E->vseg, &E->ids);
  E->vgrid = vs.grid;
  for (Chain& c : E->chains) c.bind(base);
  E->have_tables = true;
  CHK(update_gcg<TC>(E));
  E->tables_dirty = false;
  E->table_sig = sig;
  E->table_nb = nb;
  E->table_ipm = ipm;
  E->tables_version += 1;
  E->cg_on_tables = 0;
  return CSIPM_OK;
}

// admit instance h (problem data z0 (device, N reals), lambda; the solution
// goes to beta_out (device, N doubles); CG iteration cap per solve, <= 0:
// 4N).  Returns the id through *id.
int engine_admit(Engine* E, csipm_op* h, const double* z0, double lam, double* beta_out,
                 int64_t cg_itmax, int* id) {
  int k = -1;
  for (int i = 0; i < E->maxinst; ++i)
    if (!E->inst[(size_t)i].h) {
      k = i;
      break;
    }
  if (k < 0) return fail(CSIPM_ERR_ALLOC, "engine full (%d instances)", E->maxinst);
  // an instance's passes run in its handle's scratch spectrum: a handle
  // serves one live instance at a time (in any engine)
  if (h->admitted.load() > 0)
    return fail(CSIPM_ERR_ARG, "operator handle is already admitted to a live instance "
                "(one handle per instance: its scratch spectrum is per handle)");
  const int64_t len = slots_of(h);
  // partial-sum capacity: 8 quantities per vector block, the CG apply's final pass
  const int64_t pcap = len / kSlotAlign * kPartPerAlign;
  const int gcg = E->tc_size == 8 ? cg_final_grid<double>(h, 1) : cg_final_grid<float>(h, 1);
  // (the apply's p'q partials and the CG update's r'z partials side by side)
  if (8 * (int64_t)vblocks(h->U) > pcap || gcg + (int64_t)vblocks(h->U) > pcap)
    return fail(CSIPM_ERR_UNSUPPORTED, "partial sums of instance exceed its share");
  int64_t off = 0;
  if (!slots_alloc(E, len, &off))
    return fail(CSIPM_ERR_ALLOC, "no free slot range of %lld", (long long)len);
  Engine::Inst& in = E->inst[(size_t)k];
  in.h = h;
  in.off = off;
  in.len = len;
  in.serial = ++E->admissions;
  in.live = true;
  in.fresh = z0 != nullptr;   // no problem data: the instance only runs CG
  in.lam = lam;
  in.z0 = z0;
  in.beta_out = beta_out;
  BInst bi{};
  bi.off = off;
  bi.n = h->U;
  bi.N = h->N;
  bi.lam = lam;
  bi.pbase = off / kSlotAlign * kPartPerAlign;
  bi.gcg = gcg;
  bi.itmax = (int)std::min<int64_t>(cg_itmax > 0 ? cg_itmax : 4 * h->N, 1 << 30);
  // the record and zeroed flags reach the device with the next record
  // upload (one staged copy for all admissions since); the slot mask is a
  // device-to-device copy in stream order
  E->info_h[(size_t)k] = bi;
  E->pending.push_back(k);
  E->records_dirty = true;
  HIPCHK(hipMemcpyAsync(E->valid + off / 32, h->valid, sizeof(uint32_t) * (size_t)((h->U + 31) / 32),
                        hipMemcpyDeviceToDevice, E->stream));
  E->active.push_back(k);
  E->tables_dirty = true;
  h->admitted.fetch_add(1);
  *id = k;
  return CSIPM_OK;
}

void engine_drop(Engine* E, int id) {
  Engine::Inst& in = E->inst[(size_t)id];
  if (in.h) in.h->admitted.fetch_sub(1);
  slots_free(E, in.off, in.len);
  in = Engine::Inst{};
  E->active.erase(std::remove(E->active.begin(), E->active.end(), id), E->active.end());
  E->tables_dirty = true;
}

int finalize(Engine* E, const BSeg& vseg, const int* ids, int nids, std::initializer_list<int> ops,
             int mode, int first, double tau, double tol, int max_iter) {
  RedSpec spec{};
  spec.nq = 0;
  for (int op : ops) spec.op[spec.nq++] = op;
  hipLaunchKernelGGL(bk_ipm_finalize, dim3(nids), dim3(kNT), 0, E->stream, vseg, ids, E->info,
                     E->flags, E->part, spec, E->scal, mode, first, tau, tol, max_iter);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// initialise the fresh instances: b = M^T z0 (u order), ||b||_inf, the starting point
int engine_init_fresh(Engine* E) {
  CHK(engine_upload_records(E));
  std::vector<int> fresh;
  for (int id : E->active)
    if (E->inst[(size_t)id].fresh) fresh.push_back(id);
  if (fresh.empty()) return CSIPM_OK;
  StatTimer tm(&E->stats[Engine::ST_INIT]);
  Chain mtz;
  BSeg seg{};
  const int* ids = nullptr;
  int grid = 0;
  int rc = CSIPM_OK;
  do {
    TableArena& A = E->init_tabs;
    A.begin();
    if ((rc = build_chain<double>(E, A, fresh, CALL_MTZ, 2, mtz)) != CSIPM_OK) break;
    VsegOff vs;
    stage_vseg(E, A, fresh, &vs);
    char* base = nullptr;
    if ((rc = A.commit(E->stream, &base)) != CSIPM_OK) break;
    mtz.bind(base);
    vs.bind(base, &seg, &ids);
    grid = vs.grid;
    if ((rc = run_chain<double>(mtz, E->stream)) != CSIPM_OK) break;
    hipLaunchKernelGGL(bk_absmax, dim3(grid), dim3(kNT), 0, E->stream, seg, E->info, E->flags,
                       E->W(V_B), E->part);
    if ((rc = finalize(E, seg, ids, (int)fresh.size(), {1}, 0, S_BNORM, 0, 0, 0)) != CSIPM_OK)
      break;
    hipLaunchKernelGGL(bk_ipm_init, dim3(grid), dim3(kNT), 0, E->stream, seg, E->info, E->flags,
                       E->W(V_BETA), E->W(V_C), E->W(V_Z1), E->W(V_Z2), E->W(V_QB), E->scal,
                       E->valid);
    if (hipGetLastError() != hipSuccess) {
      rc = fail(CSIPM_ERR_HIP, "init launch failed");
      break;
    }
  } while (false);
  mtz.release();
  if (rc) return rc;
  for (int id : fresh) E->inst[(size_t)id].fresh = false;
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// batched preconditioned CG on the engine's CG vectors: b = C2_RHS,
// x = C2_D, diagonals C_L1.. (nb = 2: condensed K with the block
// preconditioner; nb = 1: reduced system with D in C_L1, P in C_P11).
// reliable > 0: true residual and restart every `reliable` iterations.
// ---------------------------------------------------------------------------
// one CG iteration (reliable: with the true-residual restart after it)
template <typename TC>
int cg_iteration(Engine* E, int nb, bool reliable) {
  const int na = (int)E->active.size();
  TC *b = E->C2<TC>(C2_RHS), *x = E->C2<TC>(C2_D), *r = E->C2<TC>(C2_R);
  TC *z = E->C2<TC>(C2_Z), *p = E->C2<TC>(C2_P), *q = E->C2<TC>(C2_Q);
  TC *P11 = E->C<TC>(C_P11), *P12 = E->C<TC>(C_P12), *P22 = E->C<TC>(C_P22);
  hipStream_t st = E->stream;
  CHK(run_chain_par<TC>(E->chains[CALL_CGA], st, E->side, E->fork_ev, E->join_ev));
  hipLaunchKernelGGL((bk_cg_update<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                     E->cg, x, r, z, p, q, P11, P12, P22, nb, E->part);
  hipLaunchKernelGGL(bk_cg_gamma, dim3(na), dim3(kNT), 0, st, E->vseg, E->ids, E->info, E->cg,
                     E->part);
  if (reliable) {
    CHK(run_chain_par<TC>(E->chains[CALL_CGR], st, E->side, E->fork_ev, E->join_ev));
    hipLaunchKernelGGL((bk_cg_residual<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                       E->cg, b, q, r, z, P11, P12, P22, nb, E->part);
    hipLaunchKernelGGL(bk_cg_regamma, dim3(na), dim3(kNT), 0, st, E->vseg, E->ids, E->info,
                       E->cg, E->part);
  }
  hipLaunchKernelGGL((bk_cg_pupdate<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                     E->cg, p, z, nb);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

constexpr int kCgChunk = 4;   // CG iterations between host checks (one graph)
// CG solves on unchanged tables before their chunks are captured as graphs:
// a capture costs about as much host time as the launches of a step's CG
// chunks (the GPU idles meanwhile), so tables that change every step (a
// serving loop admitting and retiring each turn) run their chunks eagerly,
// and a batch whose instance set holds for a whole step replays graphs.  A
// small batch (under kGraphAnyBlocks vector blocks of 1024 slots) is
// launch-bound: its chunks are captured at once.
constexpr int kGraphAfter = 3;
constexpr int kGraphAnyBlocks = 4096;

// capture the two CG chunks for the current tables
template <typename TC>
int capture_cg_graphs(Engine* E, int nb, int reliable) {
  StatTimer tm(&E->stats[Engine::ST_CAPTURE]);
  E->stats[Engine::ST_NCAPTURE] += 1;
  E->release_graphs();
  std::lock_guard<std::mutex> lock(g_capture_mu);
  for (int g = 0; g < (reliable > 0 ? 2 : 1); ++g) {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeRelaxed));
    int rc = CSIPM_OK;
    for (int k = 0; k < kCgChunk && rc == CSIPM_OK; ++k)
      rc = cg_iteration<TC>(E, nb, g == 1 && k == kCgChunk - 1);
    const hipError_t ec = hipStreamEndCapture(E->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ec != hipSuccess)
      return fail(CSIPM_ERR_HIP, "CG graph capture failed: %s", hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&E->graph[g], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
      return fail(CSIPM_ERR_HIP, "CG graph instantiation failed: %s", hipGetErrorString(ei));
  }
  E->graph_nb = nb;
  E->graph_rel = reliable;
  E->graph_version = E->tables_version;
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// batched preconditioned CG on the engine's CG vectors: b = C2_RHS,
// x = C2_D, diagonals C_L1.. (nb = 2: condensed K with the block
// preconditioner; nb = 1: reduced system with D in C_L1, P in C_P11).
// reliable > 0 (a multiple of kCgChunk): true residual and restart every
// `reliable` iterations.  The host checks the instances' states every
// kCgChunk iterations; the chunks in between are replayed HIP graphs.
// ---------------------------------------------------------------------------
template <typename TC>
int engine_cg(Engine* E, int nb, double atol, double rtol, int reliable, bool warm = false) {
  StatTimer tm(&E->stats[Engine::ST_CG]);
  const int na = (int)E->active.size();
  TC *b = E->C2<TC>(C2_RHS), *x = E->C2<TC>(C2_D), *r = E->C2<TC>(C2_R);
  TC *z = E->C2<TC>(C2_Z), *p = E->C2<TC>(C2_P);
  TC *P11 = E->C<TC>(C_P11), *P12 = E->C<TC>(C_P12), *P22 = E->C<TC>(C_P22);
  hipStream_t st = E->stream;
  if (warm)
    hipLaunchKernelGGL((bk_cg_init_warm<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                       E->flags, b, r, z, p, P11, E->part);
  else
    hipLaunchKernelGGL((bk_cg_init<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                       E->flags, b, x, r, z, p, P11, P12, P22, nb, E->part);
  hipLaunchKernelGGL(bk_cg_start, dim3(na), dim3(kNT), 0, st, E->vseg, E->ids, E->info, E->flags,
                     E->cg, E->part, atol, rtol, warm ? 1 : 0);
  HIPCHK(hipGetLastError());
  bool graphs = E->use_graphs && (E->cg_on_tables >= kGraphAfter || E->vgrid < kGraphAnyBlocks);
  E->cg_on_tables += 1;
  if (graphs && (E->graph_version != E->tables_version || E->graph_nb != nb ||
                 E->graph_rel != reliable))
    CHK(capture_cg_graphs<TC>(E, nb, reliable));
  for (int64_t chunk = 0;; ++chunk) {
    HIPCHK(hipMemcpyAsync(E->cg_h, E->cg, sizeof(CgState) * (size_t)E->maxinst,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    bool all = true;
    for (int id : E->active) all = all && E->cg_h[id].done;
    if (all) break;
    const bool rel = reliable > 0 && ((chunk + 1) * kCgChunk) % reliable == 0;
    if (graphs) {
      HIPCHK(hipGraphLaunch(E->graph[rel ? 1 : 0], st));
    } else {
      for (int k = 0; k < kCgChunk; ++k)
        CHK(cg_iteration<TC>(E, nb, rel && k == kCgChunk - 1));
    }
  }
  hipLaunchKernelGGL(bk_cg_account, dim3((na + 255) / 256), dim3(256), 0, st, na, E->ids, E->cg,
                     E->flags);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// one Newton solve.  FP64: CG on the condensed system (the reference's K),
// direction in C2_D.  Mixed: the c-block is eliminated in FP64
// (dc = (h_c - L2 db) / L1 leaves (M^T M + D) db = h_b - (L2/L1) h_c with
// D = L1 - L2^2/L1 = 4 S1 S2 / (S1 + S2) > 0) and the reduced system is
// solved by FP32 CG with the diagonal preconditioner 1/(1 + D); the final
// direction of each step (the corrector, refine = true) gets one FP64
// refinement step: residual in FP64, correction by FP32 CG to kRefineRtol of
// that residual.  db: the predictor's in C2_D (TC), the refined corrector's in
// V2_DD (FP64); the direction kernel back-substitutes dc.  In FP32 the
// condensed K itself is unusable near the solution: L1 and |L2| agree to
// more than 24 bits there.
// ---------------------------------------------------------------------------
constexpr double kRefineRtol = 1e-2;
// the corrector's CG starts from the predictor's solution, whose residual
// for the corrector's system the right-hand side kernel forms
constexpr bool kWarmCorrector = true;

template <typename TC>
int engine_newton(Engine* E, double atol, double rtol, int reliable, bool refine) {
  if (!E->mixed) return engine_cg<TC>(E, 2, atol, rtol, 0);
  hipStream_t st = E->stream;
  const BSeg& sg = E->vseg;
  CHK(engine_cg<TC>(E, 1, 0.0, rtol, reliable, refine && kWarmCorrector));
  if (!refine) return CSIPM_OK;
  hipLaunchKernelGGL((bk_convert<TC, double>), dim3(E->vgrid), dim3(kNT), 0, st, sg, E->info,
                     E->flags, E->C2<TC>(C2_D), E->W2(V2_DD), 1);
  hipLaunchKernelGGL((bk_convert<double, double>), dim3(E->vgrid), dim3(kNT), 0, st, sg, E->info,
                     E->flags, E->W2(V2_RHSD), E->W2(V2_TMP), 1);
  CHK(run_chain_par<double>(E->chains[CALL_REF], st, E->side, E->fork_ev, E->join_ev));
  hipLaunchKernelGGL((bk_convert<double, TC>), dim3(E->vgrid), dim3(kNT), 0, st, sg, E->info,
                     E->flags, E->W2(V2_TMP), E->C2<TC>(C2_RHS), 1);
  HIPCHK(hipGetLastError());
  CHK(engine_cg<TC>(E, 1, 0.0, kRefineRtol, reliable));
  hipLaunchKernelGGL((bk_refine_add<TC>), dim3(E->vgrid), dim3(kNT), 0, st, sg, E->info,
                     E->flags, E->W2(V2_DD), E->C2<TC>(C2_D), E->W2(V2_TMP),
                     E->C2<TC>(C2_RHS), E->C2<TC>(C2_R), E->C<TC>(C_L1));
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// One interior-point iteration of every active instance: residuals and the
// stopping test, then (for those still running) a Mehrotra predictor-
// corrector step.  *running = number of instances still iterating.
// ---------------------------------------------------------------------------
template <typename TC>
int engine_step(Engine* E, double tol, int max_iter, double cg_atol, double cg_rtol,
                int* running) {
  StatTimer tm(&E->stats[Engine::ST_STEP]);
  E->stats[Engine::ST_NSTEP] += 1;
  const int nb = E->mixed ? 1 : 2;
  CHK(engine_init_fresh(E));
  CHK(engine_tables<TC>(E, nb, true));
  const int na = (int)E->active.size();
  *running = 0;
  if (na == 0) return CSIPM_OK;
  hipStream_t st = E->stream;
  const BSeg& sg = E->vseg;
  const int g = E->vgrid;
  const double tau = 0.995;
  // FP64: the caller's PCG rule sqrt(r'Pr) <= cg_atol + cg_rtol sqrt(r0'P r0)
  // (the reference passes atol = 1e-12, rtol = 0).  Mixed: the FP32 CG stops
  // at a relative 1e-5 at the tightest (an absolute FP64-level target is out
  // of FP32's reach); accuracy comes from the FP64 refinement
  const double rtol = E->mixed ? std::max(cg_rtol, 1e-5) : cg_rtol;
  const double atol = E->mixed ? 0.0 : cg_atol;
  const int reliable = E->mixed ? 6 * kCgChunk : 0;
  double *b = E->W(V_B), *beta = E->W(V_BETA), *c = E->W(V_C), *z1 = E->W(V_Z1);
  double *z2 = E->W(V_Z2), *Qb = E->W(V_QB), *s1 = E->W(V_S1), *s2 = E->W(V_S2);
  double *dt1 = E->W(V_DT1), *dt2 = E->W(V_DT2), *dz1 = E->W(V_DZ1), *dz2 = E->W(V_DZ2);
  // Q beta = M^T M beta: an FP64 apply per step in FP64; mixed, it is carried
  // from beta = 0 by the updates beta += alpha d, Q beta += alpha M^T M d,
  // M^T M d from the corrector's FP64 refinement apply (its uncorrected
  // part) and its refinement CG (the correction's part, b - r - D x in FP32:
  // the correction is a 1e-5 part of d, so FP64-level accuracy)
  TC* rhs = E->C2<TC>(C2_RHS);
  double* rhsd = E->mixed ? E->W2(V2_RHSD) : nullptr;
  if (!E->mixed) CHK(run_chain<double>(E->chains[CALL_QB], st));
  // mixed: the reduced system comes from the residual kernel's own loads
  if (E->mixed)
    hipLaunchKernelGGL((bk_ipm_resid_prep_red<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info,
                       E->flags, beta, c, z1, z2, Qb, b, E->part, s1, s2, E->C<TC>(C_L1),
                       E->C<TC>(C_P11), rhs, E->W(V_DGD), rhsd, E->valid);
  else
    hipLaunchKernelGGL(bk_ipm_resid, dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags, beta, c,
                       z1, z2, Qb, b, E->part, E->valid);
  CHK(finalize(E, sg, E->ids, na, {1, 1, 0, 0, 0, 0}, 0, S_RB, tau, tol, max_iter));
  HIPCHK(hipMemcpyAsync(E->flags_h, E->flags, sizeof(BFlags) * (size_t)E->maxinst,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  int run = 0;
  for (int id : E->active) run += E->flags_h[id].done ? 0 : 1;
  *running = run;
  if (run == 0) return CSIPM_OK;
  if (!E->mixed)
    hipLaunchKernelGGL((bk_ipm_prep<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags, beta,
                       c, z1, z2, Qb, b, s1, s2, E->C<TC>(C_L1), E->C<TC>(C_L2),
                       E->C<TC>(C_P11), E->C<TC>(C_P12), E->C<TC>(C_P22), rhs, E->valid);
  HIPCHK(hipGetLastError());
  // predictor
  CHK(engine_newton<TC>(E, atol, rtol, reliable, false));
  auto dir = [&](int corr) -> int {
    if (E->mixed && corr)
      hipLaunchKernelGGL((bk_ipm_dir_red<double>), dim3(g), dim3(kNT), 0, st, sg, E->info,
                         E->flags, beta, c, z1, z2, s1, s2, E->W2(V2_DD), rhsd, E->scal, 1, dt1,
                         dt2, dz1, dz2, E->W2(V2_DD), E->part, E->valid);
    else if (E->mixed)
      hipLaunchKernelGGL((bk_ipm_dir_red<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags,
                         beta, c, z1, z2, s1, s2, E->C2<TC>(C2_D), rhsd, E->scal, 0, dt1, dt2,
                         dz1, dz2, nullptr, E->part, E->valid);
    else
      hipLaunchKernelGGL((bk_ipm_dir<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags, beta,
                         c, z1, z2, s1, s2, E->C2<TC>(C2_D), E->scal, corr, dt1, dt2, dz1, dz2,
                         E->part, E->valid);
    HIPCHK(hipGetLastError());
    return CSIPM_OK;
  };
  CHK(dir(0));
  if (E->mixed) {
    // step lengths and sigma in one pass (the direction kernel's sums)
    CHK(finalize(E, sg, E->ids, na, {2, 2, 0, 0, 0, 0}, 4, 0, tau, tol, max_iter));
  } else {
    CHK(finalize(E, sg, E->ids, na, {2, 2}, 1, 0, tau, tol, max_iter));
    hipLaunchKernelGGL(bk_ipm_muaff, dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags, beta, c,
                       z1, z2, dt1, dt2, dz1, dz2, E->scal, E->part, E->valid);
    CHK(finalize(E, sg, E->ids, na, {0}, 2, 0, tau, tol, max_iter));
  }
  // corrector
  if (E->mixed)
    hipLaunchKernelGGL((bk_ipm_rhs_corr_red<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info,
                       E->flags, beta, c, s1, s2, dt1, dt2, dz1, dz2, E->scal, rhs, rhsd,
                       kWarmCorrector ? E->C2<TC>(C2_R) : nullptr, E->valid);
  else
    hipLaunchKernelGGL((bk_ipm_rhs_corr<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags,
                       beta, c, z1, z2, Qb, b, dt1, dt2, dz1, dz2, E->scal, rhs, E->valid);
  HIPCHK(hipGetLastError());
  CHK(engine_newton<TC>(E, atol, rtol, reliable, true));
  CHK(dir(1));
  CHK(finalize(E, sg, E->ids, na, {2, 2}, 3, 0, tau, tol, max_iter));
  if (E->mixed)
    hipLaunchKernelGGL((bk_ipm_update<double>), dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags,
                       beta, c, z1, z2, E->W2(V2_DD), dz1, dz2, Qb, E->W2(V2_TMP), E->scal,
                       E->valid);
  else
    hipLaunchKernelGGL((bk_ipm_update<TC>), dim3(g), dim3(kNT), 0, st, sg, E->info, E->flags,
                       beta, c, z1, z2, E->C2<TC>(C2_D), dz1, dz2, nullptr, nullptr, E->scal,
                       E->valid);
  HIPCHK(hipGetLastError());
  return CSIPM_OK;
}

// results of a finished instance: beta (reference layout) into its beta_out,
// info[8] = [objective, IPM iterations, CG iterations, status, dual residual,
// mu, 0, 0]; the instance is retired and its slots freed
// sync = false: the flags the last step (or poll) read are used when they
// show the instance done (its flags no longer change), and the conversion of
// beta is left in flight on the engine's stream
int engine_harvest(Engine* E, int id, double* info, bool sync = true) {
  Engine::Inst& in = E->inst[(size_t)id];
  CHK(from_u<double>(in.h, E->W(V_BETA) + in.off, in.beta_out, E->stream));
  if (in.c_out) CHK(from_u<double>(in.h, E->W(V_C) + in.off, in.c_out, E->stream, 0));
  if (sync || !E->flags_h[id].done) {
    HIPCHK(hipMemcpyAsync(E->flags_h + id, E->flags + id, sizeof(BFlags), hipMemcpyDeviceToHost,
                          E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
  }
  const BFlags& f = E->flags_h[id];
  if (info) {
    info[0] = f.f;
    info[1] = f.iters;
    info[2] = f.cgits;
    info[3] = f.status;
    info[4] = f.dual;
    info[5] = f.mu;
    info[6] = 0;
    info[7] = 0;
  }
  engine_drop(E, id);
  return CSIPM_OK;
}

template <typename TC>
int engine_run(Engine* E, double tol, int max_iter, double cg_atol, double cg_rtol) {
  for (;;) {
    int running = 0;
    CHK(engine_step<TC>(E, tol, max_iter, cg_atol, cg_rtol, &running));
    if (running == 0) return CSIPM_OK;
  }
}

// per-sub timing of the CG apply chain of the active set (p = 0 -> q), HIP
// events around every launch, averaged over nrep applies; bytes: the
// algorithmic HBM bytes of each launch (sum over its instances)
template <typename TC>
int engine_apply_timed(Engine* E, int64_t nrep, double* ms, double* bytes, std::string* names,
                       int cap, int* nsub) {
  const int nb = E->mixed ? 1 : 2;
  CHK(engine_init_fresh(E));
  CHK(engine_tables<TC>(E, nb, true));
  const Chain& c = E->chains[CALL_CGA];
  const int ns = (int)c.subs.size();
  if (ns > cap) return fail(CSIPM_ERR_ARG, "%d launches exceed the output arrays", ns);
  HIPCHK(hipMemsetAsync(E->C2<TC>(C2_P), 0, sizeof(TC) * 2 * (size_t)E->cap, E->stream));
  // every instance runs: mark the CG states not done
  std::vector<CgState> live((size_t)E->maxinst);
  for (auto& cs : live) cs = CgState{};
  HIPCHK(hipMemcpyAsync(E->cg, live.data(), sizeof(CgState) * live.size(), hipMemcpyHostToDevice,
                        E->stream));
  std::vector<hipEvent_t> ev((size_t)ns + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  std::vector<double> acc((size_t)ns, 0.0);
  for (int64_t r = 0; r < nrep; ++r) {
    HIPCHK(hipEventRecord(ev[0], E->stream));
    for (int k = 0; k < ns; ++k) {
      const ChainSub& sb = c.subs[(size_t)k];
      BSeg seg{sb.seg, sb.seg + sb.grid};
      if (sb.nt == kNT) CHK((launch_spec<TC, kNT>(sb.spec, sb.grid, sb.lds, E->stream, sb.tab, seg)));
      else if (sb.nt == kNTSmall)
        CHK((launch_spec<TC, kNTSmall>(sb.spec, sb.grid, sb.lds, E->stream, sb.tab, seg)));
      else CHK((launch_spec<TC, kNTBig>(sb.spec, sb.grid, sb.lds, E->stream, sb.tab, seg)));
      HIPCHK(hipEventRecord(ev[(size_t)k + 1], E->stream));
    }
    HIPCHK(hipEventSynchronize(ev[(size_t)ns]));
    for (int k = 0; k < ns; ++k) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, ev[(size_t)k], ev[(size_t)k + 1]));
      acc[(size_t)k] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  const int layout = E->mixed ? 2 : 1;
  for (int k = 0; k < ns; ++k) {
    const ChainSub& sb = c.subs[(size_t)k];
    double b = 0;
    const char* nm[8];
    double by[8];
    for (int id : sb.ids) {
      kkt_pass_info(E->inst[(size_t)id].h, layout, nm, by);
      b += by[sb.pass];
    }
    kkt_pass_info(E->inst[(size_t)sb.ids[0]].h, layout, nm, by);
    ms[k] = acc[(size_t)k] / (double)nrep;
    bytes[k] = b;
    names[k] = std::to_string(sb.spec.D) + (sb.spec.kind >= PK_COL ? "d-4step:" : "d:") +
               nm[sb.pass];
    // a pass split over several launches: each name carries its workgroup
    // size and LDS bytes
    int same = 0;
    for (const ChainSub& o : c.subs)
      if (o.pass == sb.pass && o.spec.D == sb.spec.D && o.spec.kind == sb.spec.kind &&
          o.spec.axis == sb.spec.axis)
        ++same;
    if (same > 1)
      names[k] += "@nt" + std::to_string(sb.nt) + "/lds" + std::to_string(sb.lds);
  }
  *nsub = ns;
  return CSIPM_OK;
}

// ---------------------------------------------------------------------------
// ADMM (src/admm.jl:8-61) on the engine's instances.  State: b0 = M^T z0
// (V_B), x0 (V_C), z (V_BETA, the result), y (V_Z1); the x-update CG runs on
// the reduced system with D = rho (C_L1) and no preconditioner (C_P11 = 1),
// Krylov.jl's cg defaults: x from 0, ||r|| <= atol + rtol ||b||, 2N iterations.
// Initial point (admm.jl:19-24): x0 = z = ones (signed into u), y = 0.
// ---------------------------------------------------------------------------
template <typename TC>
int engine_admm_init_fresh(Engine* E, double rho) {
  CHK(engine_upload_records(E));
  std::vector<int> fresh;
  for (int id : E->active)
    if (E->inst[(size_t)id].fresh) fresh.push_back(id);
  if (fresh.empty()) return CSIPM_OK;
  Chain mtz;
  BSeg seg{};
  const int* ids = nullptr;
  int grid = 0;
  int rc = CSIPM_OK;
  do {
    TableArena& A = E->init_tabs;
    A.begin();
    if ((rc = build_chain<double>(E, A, fresh, CALL_MTZ, 1, mtz))) break;
    VsegOff vs;
    stage_vseg(E, A, fresh, &vs);
    char* base = nullptr;
    if ((rc = A.commit(E->stream, &base))) break;
    mtz.bind(base);
    vs.bind(base, &seg, &ids);
    grid = vs.grid;
    if ((rc = run_chain<double>(mtz, E->stream))) break;
    for (int id : fresh) {
      Engine::Inst& in = E->inst[(size_t)id];
      double* ones = E->W(V_DT1) + in.off;   // beta-layout scratch (N <= U slots)
      hipLaunchKernelGGL(k_fill, dim3(grid_for(in.h->N, 4)), dim3(kNT), 0, E->stream, ones,
                         in.h->N, 1.0);
      if ((rc = to_u<double>(in.h, ones, E->W(V_C) + in.off, E->stream))) break;
    }
    if (rc) break;
    hipLaunchKernelGGL((bk_admm_setup<TC>), dim3(grid), dim3(kNT), 0, E->stream, seg, E->info,
                       E->flags, E->W(V_C), E->W(V_BETA), E->W(V_Z1), E->C<TC>(C_L1),
                       E->C<TC>(C_P11), rho, E->valid);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(E->stream) != hipSuccess)
      rc = fail(CSIPM_ERR_HIP, "ADMM initialisation failed");
  } while (false);
  mtz.release();
  if (rc) return rc;
  for (int id : fresh) E->inst[(size_t)id].fresh = false;
  return CSIPM_OK;
}

// one ADMM iteration of every running instance; *running: those still iterating
template <typename TC>
int engine_admm_step(Engine* E, double rho, int maxt, double tol, double cg_atol,
                     double cg_rtol, int* running) {
  StatTimer tm(&E->stats[Engine::ST_STEP]);
  E->stats[Engine::ST_NSTEP] += 1;
  CHK(engine_admm_init_fresh<TC>(E, rho));
  CHK(engine_tables<TC>(E, 1, false));
  const int na = (int)E->active.size();
  *running = 0;
  if (na == 0) return CSIPM_OK;
  hipStream_t st = E->stream;
  hipLaunchKernelGGL((bk_admm_rhs<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                     E->flags, E->W(V_B), E->W(V_BETA), E->W(V_Z1), rho, E->C2<TC>(C2_RHS),
                     E->valid);
  HIPCHK(hipGetLastError());
  CHK(engine_cg<TC>(E, 1, cg_atol, cg_rtol, 0));
  hipLaunchKernelGGL((bk_admm_update<TC>), dim3(E->vgrid), dim3(kNT), 0, st, E->vseg, E->info,
                     E->flags, E->C2<TC>(C2_D), E->W(V_C), E->W(V_BETA), E->W(V_Z1), rho,
                     E->part, E->valid);
  hipLaunchKernelGGL(bk_admm_finalize, dim3(na), dim3(kNT), 0, st, E->vseg, E->ids, E->info,
                     E->flags, E->part, tol, maxt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(E->flags_h, E->flags, sizeof(BFlags) * (size_t)E->maxinst,
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int id : E->active) *running += E->flags_h[id].done ? 0 : 1;
  return CSIPM_OK;
}
