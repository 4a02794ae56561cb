// The batched form: the blocks of many files, of uneven lengths, planned together (included by capi.hip).
//
// A block of r bytes runs in a slab of 2^c bytes, c = max(10, ceil(log2 r)) up to the archive's block_log (its size
// class): the position-parallel stages then do work for 2^c positions instead of 2^block_log, and the per-block
// scratch shrinks the same way. The hash keys keep the archive's width, so every block's payload is the one the
// plain path writes.
//
// Every stage is one persistent kernel over the blocks of every class: a grid of a few workgroups per compute unit
// that take work items (a block, a stream of a block, a partition, 256 sorted slots, ...) from device queues until
// they run dry, and run on each item the same code as the plain path's kernel for it. Items come largest class
// first, so the long blocks start early and the short ones fill the end. There are eight queues per stage, one per
// XCD (a workgroup serves the queue of the XCD it runs on, workgroup i running on XCD i % 8): the items of a block
// stay on one XCD, so the block's data and scattered results stay in one L2. Encoder scratch is about 19 bytes per
// slab byte, so the blocks run in passes of at most PASS_SLAB slab bytes, one scratch arena reused pass after pass;
// a pass is the stage sequence over its blocks, each stage one launch.
//
// The launch sequence of a call can be captured into a hipGraph the first time and replayed when the same buffers
// come again. zgpu_plan_tune chooses the output-invariant parameters: each class's bucket bits (the hash-bucket
// depth) and parse walkers (the parse-group width), and each stage's workgroups per compute unit.
#include <algorithm>
#include <map>
#include <tuple>

namespace {

__global__ void k_scatter_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                              const uint32_t* __restrict__ order, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[order[i]] = src[i];
}

enum PStage : uint32_t {
  P_HASH, P_PLACE, P_SORT_S, P_SORT_M, P_SORT_L, P_MATCH, P_PARSE, P_TOKENS, P_RANS, P_PACK, P_DSTREAMS, P_DLZ,
  P_STAGES
};
constexpr const char* PSTAGE_NAME[P_STAGES] = {"hash",   "place",  "sort_s", "sort_m", "sort_l",   "match",
                                               "parse",  "tokens", "rans",   "pack",   "dstreams", "dlz"};
constexpr uint32_t MAX_SEGS = 16;            // classes (or parts of one) in a pass
constexpr uint64_t PASS_SLAB = 1ull << 31;   // slab bytes of a pass (by default)
constexpr uint32_t QSTRIDE = 32;             // u32s from one queue head to the next: a 128-byte line each
constexpr size_t CTR_WORDS = P_STAGES * 8 * QSTRIDE + 4 * MAX_SEGS;

// one class's blocks in a pass, as the persistent kernels see them (device memory)
struct PSeg {
  uint32_t first, count;  // plan-order blocks
  uint32_t blog, hbits, pbits, keybits, walkers, pad;
  uint64_t slab_off;      // the first block's slab, from the slab base (the others follow, 2^blog apart)
  uint32_t *srt, *mt, *bstart, *mark;
  SortQueues q;
  TokBufs T;
  RansBufs R;
  uint8_t *dlit, *dtok, *dext, *dsym, *dgrp;  // decoder scratch
};
struct PTab {
  uint32_t nseg, pad;
  PSeg s[MAX_SEGS];  // largest slab first
};

// item j of XCD queue x, for a stage with per(seg) items per block: its segment, block in the segment and item in
// the block; false past the last item. The blocks of a segment go to the queues in turn (block k to queue k % 8).
template <class Per>
__device__ inline bool locate(const PTab* __restrict__ t, uint32_t x, uint32_t j, Per per, uint32_t& c, uint32_t& bl,
                              uint32_t& sub) {
  for (c = 0; c < t->nseg; ++c) {
    const uint32_t cnt = t->s[c].count;
    const uint32_t nbx = cnt > x ? (cnt - x + 7u) >> 3 : 0u;
    const uint32_t pc = per(t->s[c]);
    const uint64_t n = (uint64_t)nbx * pc;
    if (j < n) {
      bl = x + 8u * (j / pc);
      sub = j % pc;
      return true;
    }
    j -= (uint32_t)n;
  }
  return false;
}
// the five streams of the blocks of queue x: every literal stream first, then the others
__device__ inline bool locate5(const PTab* __restrict__ t, uint32_t x, uint32_t j, uint32_t& c, uint32_t& bl,
                               uint32_t& s) {
  uint32_t tot = 0;
  for (uint32_t k = 0; k < t->nseg; ++k) tot += t->s[k].count > x ? (t->s[k].count - x + 7u) >> 3 : 0u;
  if (tot == 0 || j >= 5u * tot) return false;
  s = j / tot;
  uint32_t sub;
  return locate(t, x, j % tot, [](const PSeg&) { return 1u; }, c, bl, sub);
}
// Claims: thread 0 takes the next n items of a queue with one atomic, and the workgroup learns the first of them.
// A workgroup claims its next items as soon as it starts on the current ones (claim_issue), and reads the answer only
// when it is done with them (claim_share): the atomic's round trip to memory, a microsecond or more, then overlaps
// the work instead of preceding every item (small blocks take a few microseconds each).
__device__ inline uint32_t claim_issue(uint32_t* q, uint32_t n) { return threadIdx.x == 0 ? atomicAdd(q, n) : 0u; }
__device__ inline uint32_t claim_share(uint32_t v) {  // thread 0's v, for every thread of the workgroup
  if (blockDim.x <= 64u) return __builtin_amdgcn_readfirstlane(v);  // (one wavefront, every lane active)
  __shared__ uint32_t got;
  __syncthreads();  // (everyone has read the last one)
  if (threadIdx.x == 0) got = v;
  __syncthreads();
  return got;
}
// the items of this workgroup's XCD queue, claimed G at a time (c, bl, sub: the item's segment, block in it, item in
// the block); the loop's exit is uniform over the workgroup, as every item is. The claim after the current one is
// in flight while the current items run (its result is unused when the queue ends first).
#define PLAN_LOOP(stage, PER, G)                                                                      \
  const uint32_t x = blockIdx.x & 7u;                                                                 \
  uint32_t c, bl, sub;                                                                                \
  uint32_t* const head = ctr + ((stage) * 8u + x) * QSTRIDE;                                          \
  uint32_t j0 = claim_share(claim_issue(head, (G))), ahead = claim_issue(head, (G));                 \
  for (uint32_t j = j0; locate(t, x, j, PER, c, bl, sub);                                             \
       j = j + 1u - j0 < (G) ? j + 1u : (j0 = claim_share(ahead), ahead = claim_issue(head, (G)), j0))

__global__ __launch_bounds__(1024) void k_plan_hash(const PTab* __restrict__ t, uint32_t* ctr, const Block* enc,
                                                    const uint8_t* in) {
  PLAN_LOOP(P_HASH, [](const PSeg&) { return 1u; }, 1u) {
    const PSeg& g = t->s[c];
    enc_hash_block(enc + g.first, in, g.blog, g.hbits, g.pbits, g.mt, g.bstart, bl);
  }
}
__global__ __launch_bounds__(1024) void k_plan_place(const PTab* __restrict__ t, uint32_t* ctr, const Block* enc,
                                                     const uint8_t* in) {
  PLAN_LOOP(P_PLACE, [](const PSeg& g) { return 1u << g.pbits; }, 4u) {
    const PSeg& g = t->s[c];
    enc_place_part(enc + g.first, in, g.blog, g.hbits, g.keybits, g.pbits, g.mt, g.bstart, g.srt, bl, sub);
  }
}
// 256 buckets of a block per item (all of them when the block has fewer)
__global__ __launch_bounds__(256) void k_plan_sort_s(const PTab* __restrict__ t, uint32_t* ctr) {
  PLAN_LOOP(P_SORT_S, [](const PSeg& g) { return g.hbits > 8u ? 1u << (g.hbits - 8u) : 1u; }, 8u) {
    const PSeg& g = t->s[c];
    const uint32_t h = sub * 256u + threadIdx.x;
    enc_sort_bucket(g.blog, g.hbits, g.srt, g.bstart, g.q, bl, h, h < (1u << g.hbits));
  }
}
__global__ __launch_bounds__(256) void k_plan_sort_m(const PTab* __restrict__ t) {
  for (uint32_t c = 0; c < t->nseg; ++c) {
    const PSeg& g = t->s[c];
    enc_sort_m_range(g.blog, g.hbits, g.srt, g.bstart, g.q, blockIdx.x, gridDim.x);
  }
}
__global__ __launch_bounds__(1024) void k_plan_sort_l(const PTab* __restrict__ t) {
  for (uint32_t c = 0; c < t->nseg; ++c) {
    const PSeg& g = t->s[c];
    enc_sort_l_range(g.blog, g.hbits, g.srt, g.mt, g.bstart, g.q, blockIdx.x, gridDim.x);
  }
}
// every segment's match array to zero (the match stage writes only the positions that get a match)
__global__ __launch_bounds__(256) void k_plan_clear(const PTab* __restrict__ t) {
  for (uint32_t c = 0; c < t->nseg; ++c) {
    const PSeg& g = t->s[c];
    uint4* m = (uint4*)g.mt;  // (256-byte aligned: every arena piece is)
    const size_t n16 = ((size_t)g.count << g.blog) / 4;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256)
      m[i] = uint4{0u, 0u, 0u, 0u};
  }
}
// an item is blockDim.x sorted slots (64..1024, the tuner's choice; blocks have 1024 or more), one wavefront per 64;
// dynamic LDS: MATCH_LDS_WORDS per wavefront
__global__ __launch_bounds__(1024) void k_plan_match(const PTab* __restrict__ t, uint32_t* ctr, const Block* enc,
                                                     const uint8_t* in, uint32_t ncand, uint32_t lencap) {
  extern __shared__ uint32_t msh[];
  const uint32_t lw = (uint32_t)__builtin_ctz(blockDim.x);
  // one item a claim: the items in flight on an XCD then cover a fraction of one block, whose scattered match writes
  // stay in that XCD's L2 until whole lines are written (16 a claim spread them over several blocks: twice the time)
  PLAN_LOOP(P_MATCH, [lw](const PSeg& g) { return 1u << (g.blog - lw); }, 1u) {
    const PSeg& g = t->s[c];
    enc_match_slots(enc + g.first, in, g.blog, ncand, lencap, g.srt, g.mt, bl, sub << lw, msh);
  }
}
__global__ __launch_bounds__(1024) void k_plan_parse(const PTab* __restrict__ t, uint32_t* ctr, const Block* enc,
                                                     const uint8_t* in, uint32_t lencap) {
  PLAN_LOOP(P_PARSE, [](const PSeg&) { return 1u; }, 1u) {
    const PSeg& g = t->s[c];
    enc_parse_block(enc + g.first, in, g.blog, g.walkers, lencap, g.mt, g.mark, bl);
  }
}
__global__ __launch_bounds__(64) void k_plan_tokens(const PTab* __restrict__ t, uint32_t* ctr, const Block* enc,
                                                    const uint8_t* in, uint32_t lencap) {
  PLAN_LOOP(P_TOKENS, [](const PSeg&) { return 1u; }, 1u) {
    const PSeg& g = t->s[c];
    enc_tokens_block(enc + g.first, in, g.blog, lencap, g.mt, g.mark, g.T, bl);
  }
}
__global__ __launch_bounds__(64) void k_plan_rans(const PTab* __restrict__ t, uint32_t* ctr) {
  const uint32_t x = blockIdx.x & 7u;
  uint32_t c, bl, s;
  uint32_t* const head = ctr + (P_RANS * 8u + x) * QSTRIDE;
  uint32_t j = claim_share(claim_issue(head, 1u)), ahead = claim_issue(head, 1u);
  for (; locate5(t, x, j, c, bl, s); j = claim_share(ahead), ahead = claim_issue(head, 1u)) {
    const PSeg& g = t->s[c];
    enc_rans_stream(g.blog, g.T, g.R, bl, s);
  }
}
__global__ __launch_bounds__(256) void k_plan_pack(const PTab* __restrict__ t, uint32_t* ctr, const Block* enc,
                                                   uint8_t* slabs, uint32_t* res) {
  PLAN_LOOP(P_PACK, [](const PSeg&) { return 1u; }, 1u) {
    const PSeg& g = t->s[c];
    enc_pack_block(enc + g.first, g.blog, g.T, g.R, slabs + g.slab_off, res + g.first, bl);
  }
}
__global__ __launch_bounds__(64) void k_plan_dstreams(const PTab* __restrict__ t, uint32_t* ctr, const Block* dec,
                                                      const uint8_t* comp, uint32_t* res) {
  const uint32_t x = blockIdx.x & 7u;
  uint32_t c, bl, s;
  uint32_t* const head = ctr + (P_DSTREAMS * 8u + x) * QSTRIDE;
  uint32_t j = claim_share(claim_issue(head, 1u)), ahead = claim_issue(head, 1u);
  for (; locate5(t, x, j, c, bl, s); j = claim_share(ahead), ahead = claim_issue(head, 1u)) {
    const PSeg& g = t->s[c];
    dec_stream(dec + g.first, comp, g.blog, g.dlit, g.dtok, g.dext, g.dsym, g.dgrp, res + g.first, bl, s);
  }
}
// (at most 96 VGPRs, so that 5 wavefronts fit a SIMD: the persistent loop's state would otherwise take it to 106)
__global__ __launch_bounds__(64, 5) void k_plan_dlz(const PTab* __restrict__ t, uint32_t* ctr, const Block* dec,
                                                 const uint8_t* comp, uint8_t* out, uint32_t* res) {
  PLAN_LOOP(P_DLZ, [](const PSeg&) { return 1u; }, 1u) {
    const PSeg& g = t->s[c];
    dec_lz_block(dec + g.first, comp, g.blog, g.dlit, g.dtok, g.dext, g.dsym, g.dgrp, out, res + g.first, bl);
  }
}
#undef PLAN_LOOP

struct PlanClass {
  EncShape shape;
  uint32_t first, count;  // plan-order block range
};
struct Seg {
  uint32_t cls, first, count;
  size_t enc_off, dec_off;  // in the arenas
};
struct Pass {
  std::vector<Seg> segs;
  size_t enc_bytes = 0, dec_bytes = 0;
  PTab* d_tab = nullptr;
};

uint32_t class_log(uint32_t raw_len, uint32_t key_log) {
  uint32_t c = 10;
  while (c < key_log && (1u << c) < raw_len) ++c;
  return c;
}
EncShape class_shape(uint32_t blog, uint32_t key_log) {
  EncShape S = default_shape(blog);
  S.keybits = hbits_of(key_log) + KEY_TAG_BITS;
  S.hbits = std::min(S.hbits, S.keybits);
  while (S.keybits - S.hbits + S.blog > 32) ++S.hbits;
  S.walkers = walkers_of(blog);
  return S;
}
size_t parse_lds_bytes(const EncShape& S) { return parse_lds_words(S.blog, S.walkers) * 4; }
size_t hash_lds_bytes(const EncShape& S) {
  return hash_lds_words(S.hbits, pbits_of(S.blog, S.hbits)) * 4;
}
size_t sort_l_lds_shape(const EncShape& S) { return sort_l_lds_bytes(S.blog); }
size_t place_lds_bytes(const EncShape& S) {
  return (((size_t)1 << (S.hbits - pbits_of(S.blog, S.hbits))) + PLACE_CAP) * 4;
}
bool shape_fits(const EncShape& S) {  // (within the plain kernels' 160 KiB and the persistent ones' 158 KiB)
  return shape_ok(S) && parse_lds_bytes(S) <= 158 * 1024 && hash_lds_bytes(S) <= 158 * 1024 &&
         place_lds_bytes(S) <= 158 * 1024;
}

// workgroups per compute unit of each stage's persistent grid, as the resources allow (LDS, registers): the tuner's
// starting point for the encoder's stages, and the decoder's setting (the LZ replay is bounded to 96 VGPRs, 5
// wavefronts per SIMD: 20 one-wavefront workgroups per CU; on 200,000 uneven files it took 22.3 ms at 12 per CU,
// 18.4 at 16 and 17.6 at 20)
constexpr uint32_t DEFAULT_OCC[P_STAGES] = {1, 2, 8, 4, 1, 8, 2, 16, 8, 8, 16, 20};

}  // namespace

struct zgpu_plan {
  zgpu_t* g = nullptr;
  uint32_t nb = 0, key_log = 18;
  std::vector<uint32_t> order;     // plan index -> caller's block index
  std::vector<uint64_t> slab_off;  // plan index -> slab offset
  uint64_t slab_total = 0;
  std::vector<PlanClass> cls;
  std::vector<Pass> passes;
  DBuf ework, dwork;  // the arenas of a pass (reused by the next)
  std::vector<Block> htab;  // plan order: src, dst = src, raw_len
  Block* d_enc = nullptr;  // plan order
  Block* d_dec = nullptr;  // plan order (after zgpu_plan_set_stored)
  uint32_t* d_order = nullptr;
  uint32_t* d_tmp = nullptr;
  uint32_t* d_ctr = nullptr;  // queue heads [P_STAGES][8] (QSTRIDE apart), then the sort queues' counts [MAX_SEGS][4]
  uint32_t occ[P_STAGES];
  uint32_t match_wg = 256;  // threads per workgroup of the match stage (tuned)
  uint64_t pass_slab = PASS_SLAB;
  bool tables_ready = false, use_graphs = false;
  std::map<std::tuple<int, const void*, const void*, const void*, int>, hipGraphExec_t> graphs;
  std::vector<hipEvent_t> stage_ev;  // (the tuner's, and the profile's)
  bool profile = false;
  int profiled_op = -1;  // the call the events belong to
};

static void plan_drop_graphs(zgpu_plan* p) {
  for (auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  p->graphs.clear();
}

static void plan_free(zgpu_plan* p) {
  plan_drop_graphs(p);
  for (DBuf* b : {&p->ework, &p->dwork}) if (b->p) (void)hipFree(b->p);
  for (Pass& ps : p->passes) if (ps.d_tab) (void)hipFree(ps.d_tab);
  for (hipEvent_t e : p->stage_ev) (void)hipEventDestroy(e);
  for (void* q : {(void*)p->d_enc, (void*)p->d_dec, (void*)p->d_order, (void*)p->d_tmp, (void*)p->d_ctr})
    if (q) (void)hipFree(q);
  delete p;
}

// the passes: classes largest first, blocks taken in order until a pass holds pass_slab slab bytes (or MAX_SEGS
// pieces of classes); the arenas' offsets of every piece
static void plan_passes(zgpu_plan* p) {
  for (Pass& ps : p->passes) if (ps.d_tab) (void)hipFree(ps.d_tab);
  p->passes.clear();
  Pass cur;
  uint64_t slab = 0;
  for (size_t ci = p->cls.size(); ci-- > 0;) {
    const PlanClass& c = p->cls[ci];
    uint32_t at = 0;
    while (at < c.count) {
      if (slab >= p->pass_slab || cur.segs.size() == MAX_SEGS) {
        p->passes.push_back(std::move(cur));
        cur = Pass();
        slab = 0;
      }
      const uint64_t room = (p->pass_slab - slab) >> c.shape.blog;
      const uint32_t k = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(room, 1), c.count - at);
      Seg s{(uint32_t)ci, c.first + at, k, cur.enc_bytes, cur.dec_bytes};
      cur.enc_bytes += al(enc_layout(k, c.shape.blog, c.shape.hbits).total);
      cur.dec_bytes += al(dec_layout(k, c.shape.blog).total);
      cur.segs.push_back(s);
      slab += (uint64_t)k << c.shape.blog;
      at += k;
    }
  }
  if (!cur.segs.empty()) p->passes.push_back(std::move(cur));
}

// arenas for the largest pass, and every pass's table with its pieces' scratch pointers
static int plan_tables(zgpu_plan* p) {
  zgpu_t* g = p->g;
  plan_passes(p);
  size_t en = 0, de = 0;
  for (const Pass& ps : p->passes) {
    en = std::max(en, ps.enc_bytes);
    de = std::max(de, ps.dec_bytes);
  }
  if (int rc = ensure(g, p->ework, std::max<size_t>(en, 256))) return rc;
  if (int rc = ensure(g, p->dwork, std::max<size_t>(de, 256))) return rc;
  for (Pass& ps : p->passes) {
    PTab tab{};
    tab.nseg = (uint32_t)ps.segs.size();
    for (size_t k = 0; k < ps.segs.size(); ++k) {
      const Seg& s = ps.segs[k];
      const EncShape& S = p->cls[s.cls].shape;
      PSeg& q = tab.s[k];
      q.first = s.first;
      q.count = s.count;
      q.blog = S.blog;
      q.hbits = S.hbits;
      q.pbits = pbits_of(S.blog, S.hbits);
      q.keybits = S.keybits;
      q.walkers = S.walkers;
      q.slab_off = p->slab_off[s.first];
      uint8_t* base = (uint8_t*)p->ework.p + s.enc_off;
      const EncLayout L = enc_layout(s.count, S.blog, S.hbits);
      q.srt = (uint32_t*)(base + L.srt);
      q.mt = (uint32_t*)(base + L.mtch);
      q.bstart = (uint32_t*)(base + L.bstart);
      q.mark = (uint32_t*)(base + L.mark);
      q.q = SortQueues{(uint32_t*)(base + L.qmid), (uint32_t*)(base + L.qlarge), p->d_ctr + P_STAGES * 8 * QSTRIDE + 4 * k, L.qcap};
      q.T = TokBufs{base + L.lit,  (uint32_t*)(base + L.LB), (uint32_t*)(base + L.TM), (uint32_t*)(base + L.TD),
                    base + L.tok,  base + L.sym,              base + L.grp,              base + L.ext,
                    (uint32_t*)(base + L.offw), (uint32_t*)(base + L.cnt)};
      q.R = RansBufs{base + L.lanes, (uint16_t*)(base + L.freq), (uint32_t*)(base + L.lanelen),
                     (uint32_t*)(base + L.sinfo)};
      uint8_t* dbase = (uint8_t*)p->dwork.p + s.dec_off;
      const DecLayout D = dec_layout(s.count, S.blog);
      q.dlit = dbase + D.lit;
      q.dtok = dbase + D.tok;
      q.dext = dbase + D.ext;
      q.dsym = dbase + D.sym;
      q.dgrp = dbase + D.grp;
    }
    if (!ps.d_tab) HIPCHK(g, hipMalloc((void**)&ps.d_tab, sizeof(PTab)));
    HIPCHK(g, hipMemcpy(ps.d_tab, &tab, sizeof tab, hipMemcpyHostToDevice));
  }
  p->tables_ready = true;
  return ZGPU_OK;
}

// the largest dynamic LDS a stage's item needs in a pass
static size_t pass_lds(const zgpu_plan* p, const Pass& ps, size_t (*f)(const EncShape&)) {
  size_t m = 0;
  for (const Seg& s : ps.segs) m = std::max(m, f(p->cls[s.cls].shape));
  return m;
}
static uint32_t pass_threads_parse(const zgpu_plan* p, const Pass& ps) {
  uint32_t m = 64;
  for (const Seg& s : ps.segs) m = std::max(m, p->cls[s.cls].shape.walkers);
  return m;
}

// a stage's persistent grid: occ workgroups per compute unit, a multiple of the 8 XCDs
static unsigned plan_grid(const zgpu_plan* p, uint32_t stage) {
  const uint32_t cus = (uint32_t)std::max(8, p->g->props.cu_count);
  return (unsigned)std::max<uint32_t>(8, (p->occ[stage] * cus + 7u) & ~7u);
}

// the launch sequence of one call: every pass, its stages one persistent launch each; then out_len / err in the
// caller's order. ev (the tuner's): an event after each stage of every pass, on the context's stream.
static int plan_launch(zgpu_plan* p, int op, const uint8_t* d_src, uint8_t* d_dst, uint32_t* d_res, int level,
                       hipEvent_t* ev) {
  zgpu_t* g = p->g;
  hipStream_t st = g->stream;
  uint32_t ncand, lencap;
  level_params(op == 0 ? (uint32_t)level : 5u, ncand, lencap);
  const size_t nctr = CTR_WORDS * 4;
  if (op == 1) k_fill<<<(unsigned)std::min<size_t>((p->nb * 4 + 255) / 256, 1024), 256, 0, st>>>((uint8_t*)p->d_tmp, 0, (size_t)p->nb * 4);
  int e = 0;
  auto mark = [&]() { if (ev) (void)hipEventRecord(ev[e++], st); };
  for (const Pass& ps : p->passes) {
    k_fill<<<1, 256, 0, st>>>((uint8_t*)p->d_ctr, 0, nctr);
    const PTab* t = ps.d_tab;
    uint32_t* q = p->d_ctr;
    if (op == 0) {
      k_plan_hash<<<plan_grid(p, P_HASH), 1024, pass_lds(p, ps, hash_lds_bytes), st>>>(t, q, p->d_enc, d_src);
      mark();
      k_plan_place<<<plan_grid(p, P_PLACE), 1024, pass_lds(p, ps, place_lds_bytes), st>>>(t, q, p->d_enc, d_src);
      mark();
      k_plan_sort_s<<<plan_grid(p, P_SORT_S), 256, 0, st>>>(t, q);
      mark();
      k_plan_sort_m<<<plan_grid(p, P_SORT_M), 256, 0, st>>>(t);
      mark();
      k_plan_sort_l<<<plan_grid(p, P_SORT_L), 1024, pass_lds(p, ps, sort_l_lds_shape), st>>>(t);
      mark();
      k_plan_clear<<<(unsigned)std::max(8, p->g->props.cu_count) * 8, 256, 0, st>>>(t);
      k_plan_match<<<plan_grid(p, P_MATCH), p->match_wg, (p->match_wg / 64) * MATCH_LDS_WORDS * 4, st>>>(
          t, q, p->d_enc, d_src, ncand, lencap);
      mark();
      k_plan_parse<<<plan_grid(p, P_PARSE), pass_threads_parse(p, ps), pass_lds(p, ps, parse_lds_bytes), st>>>(
          t, q, p->d_enc, d_src, lencap);
      mark();
      k_plan_tokens<<<plan_grid(p, P_TOKENS), 64, 0, st>>>(t, q, p->d_enc, d_src, lencap);
      mark();
      k_plan_rans<<<plan_grid(p, P_RANS), 64, 0, st>>>(t, q);
      mark();
      k_plan_pack<<<plan_grid(p, P_PACK), 256, 0, st>>>(t, q, p->d_enc, d_dst, p->d_tmp);
      mark();
    } else {
      k_plan_dstreams<<<plan_grid(p, P_DSTREAMS), 64, 0, st>>>(t, q, p->d_dec, d_src, p->d_tmp);
      mark();
      k_plan_dlz<<<plan_grid(p, P_DLZ), 64, 0, st>>>(t, q, p->d_dec, d_src, d_dst, p->d_tmp);
      mark();
    }
    HIPCHK(g, hipGetLastError());
  }
  k_scatter_u32<<<(p->nb + 255) / 256, 256, 0, st>>>(d_res, p->d_tmp, p->d_order, p->nb);
  HIPCHK(g, hipGetLastError());
  return ZGPU_OK;
}

static int plan_events(zgpu_plan* p) {  // one before the first stage, one after each stage of every pass
  const size_t nev = 1 + p->passes.size() * 10;
  while (p->stage_ev.size() < nev) {
    hipEvent_t e;
    HIPCHK(p->g, hipEventCreate(&e));
    p->stage_ev.push_back(e);
  }
  return ZGPU_OK;
}

static int plan_call(zgpu_plan* p, int op, const uint8_t* d_src, uint8_t* d_dst, uint32_t* d_res, int level) {
  if (p->nb == 0) return ZGPU_OK;
  if (!p->tables_ready)
    if (int rc = plan_tables(p)) return rc;
  if (p->profile) {  // (events between the stages: no graph)
    if (int rc = plan_events(p)) return rc;
    HIPCHK(p->g, hipEventRecord(p->stage_ev[0], p->g->stream));
    p->profiled_op = op;
    return plan_launch(p, op, d_src, d_dst, d_res, level, p->stage_ev.data() + 1);
  }
  if (!p->use_graphs) return plan_launch(p, op, d_src, d_dst, d_res, level, nullptr);
  const auto key = std::make_tuple(op, (const void*)d_src, (const void*)d_dst, (const void*)d_res, level);
  auto it = p->graphs.find(key);
  if (it == p->graphs.end()) {
    zgpu_t* g = p->g;
    hipGraph_t graph;
    HIPCHK(g, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    const int rc = plan_launch(p, op, d_src, d_dst, d_res, level, nullptr);
    const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
    if (rc) return rc;
    HIPCHK(g, ec);
    hipGraphExec_t exec;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(g, ei);
    it = p->graphs.emplace(key, exec).first;
  }
  HIPCHK(p->g, hipGraphLaunch(it->second, p->g->stream));
  return ZGPU_OK;
}

extern "C" {

int zgpu_plan_create(zgpu_t* g, uint32_t nb, const uint64_t* src, const uint32_t* raw_len, uint32_t key_log,
                     int streams, zgpu_plan_t** out) {
  *out = nullptr;
  if (key_log < 16 || key_log > 20) return fail(g, ZGPU_ERR_ARG, "block_log 16..20");
  if (streams < 1 || streams > 8) return fail(g, ZGPU_ERR_ARG, "streams 1..8");
  for (uint32_t b = 0; b < nb; ++b)
    if (raw_len[b] > (1u << key_log)) return fail(g, ZGPU_ERR_ARG, "a block is longer than 2^block_log");
  // (dynamic LDS up to 158 KiB: the kernels' static LDS, the queue slot, must fit beside it in the 160 KiB)
  for (const void* k : {reinterpret_cast<const void*>(&k_plan_hash), reinterpret_cast<const void*>(&k_plan_place),
                        reinterpret_cast<const void*>(&k_plan_parse)})
    HIPCHK(g, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 158 * 1024));
  HIPCHK(g, hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plan_sort_l),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sort_l_lds_bytes(20)));
  zgpu_plan* p = new zgpu_plan();
  p->g = g;
  p->nb = nb;
  p->key_log = key_log;
  std::copy(DEFAULT_OCC, DEFAULT_OCC + P_STAGES, p->occ);
  // classes in increasing slab size; blocks keep their relative order inside a class
  std::vector<uint32_t> cl(nb);
  for (uint32_t b = 0; b < nb; ++b) cl[b] = class_log(raw_len[b], key_log);
  p->order.resize(nb);
  for (uint32_t b = 0; b < nb; ++b) p->order[b] = b;
  std::stable_sort(p->order.begin(), p->order.end(), [&](uint32_t a, uint32_t b) { return cl[a] < cl[b]; });
  p->slab_off.resize(nb);
  for (uint32_t i = 0; i < nb;) {
    const uint32_t c = cl[p->order[i]];
    uint32_t j = i;
    while (j < nb && cl[p->order[j]] == c) ++j;
    p->cls.push_back(PlanClass{class_shape(c, key_log), i, j - i});
    for (uint32_t k = i; k < j; ++k) {
      p->slab_off[k] = p->slab_total;
      p->slab_total += (uint64_t)1 << c;
    }
    i = j;
  }
  int rc = ZGPU_OK;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == ZGPU_OK) rc = hip_fail(g, e, what);
  };
  chk(hipMalloc((void**)&p->d_ctr, CTR_WORDS * 4), "hipMalloc");
  if (nb) {
    std::vector<Block>& tab = p->htab;
    tab.resize(nb);
    for (uint32_t i = 0; i < nb; ++i) {
      const uint32_t b = p->order[i];
      tab[i] = Block{src[b], src[b], raw_len[b], 0, 0, 0};
    }
    chk(hipMalloc((void**)&p->d_enc, nb * sizeof(Block)), "hipMalloc");
    chk(hipMalloc((void**)&p->d_order, nb * 4), "hipMalloc");
    chk(hipMalloc((void**)&p->d_tmp, nb * 4), "hipMalloc");
    if (rc == ZGPU_OK) chk(hipMemcpy(p->d_enc, tab.data(), nb * sizeof(Block), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc == ZGPU_OK) chk(hipMemcpy(p->d_order, p->order.data(), nb * 4, hipMemcpyHostToDevice), "hipMemcpy");
  }
  (void)streams;  // (one stream: the passes reuse one scratch arena; every stage spreads over the whole GPU itself)
  if (rc != ZGPU_OK) { plan_free(p); return rc; }
  *out = p;
  return ZGPU_OK;
}

void zgpu_plan_destroy(zgpu_plan_t* p) {
  if (p) {
    (void)hipStreamSynchronize(p->g->stream);
    plan_free(p);
  }
}

uint64_t zgpu_plan_slab_bytes(const zgpu_plan_t* p) { return p->slab_total; }

int zgpu_plan_slab_offsets(const zgpu_plan_t* p, uint64_t* slab_off) {
  for (uint32_t i = 0; i < p->nb; ++i) slab_off[p->order[i]] = p->slab_off[i];
  return ZGPU_OK;
}

int zgpu_plan_classes(const zgpu_plan_t* p, uint32_t* block_log, uint32_t* blocks, uint32_t* hash_bits,
                      uint32_t* walkers, int cap) {
  int k = 0;
  for (const PlanClass& c : p->cls) {
    if (k >= cap) break;
    block_log[k] = c.shape.blog;
    blocks[k] = c.count;
    hash_bits[k] = c.shape.hbits;
    walkers[k] = c.shape.walkers;
    ++k;
  }
  return (int)p->cls.size();
}

int zgpu_plan_stage_occupancy(zgpu_plan_t* p, uint32_t* occ, int cap) {
  for (int k = 0; k < cap && k < (int)P_STAGES; ++k) occ[k] = p->occ[k];
  return (int)P_STAGES;
}

int zgpu_plan_set_stage_occupancy(zgpu_plan_t* p, const uint32_t* occ, int n) {
  for (int k = 0; k < n && k < (int)P_STAGES; ++k)
    if (occ[k] < 1 || occ[k] > 64) return fail(p->g, ZGPU_ERR_ARG, "1..64 workgroups per compute unit");
  for (int k = 0; k < n && k < (int)P_STAGES; ++k) p->occ[k] = occ[k];
  plan_drop_graphs(p);
  return ZGPU_OK;
}

int zgpu_plan_match_workgroup(zgpu_plan_t* p, uint32_t threads) {
  if (threads) {
    if (threads < 64 || threads > 1024 || (threads & (threads - 1)))
      return fail(p->g, ZGPU_ERR_ARG, "match workgroup: 64, 128, 256, 512 or 1024 threads");
    p->match_wg = threads;
    plan_drop_graphs(p);
  }
  return (int)p->match_wg;
}

int zgpu_plan_pass_bytes(zgpu_plan_t* p, uint64_t slab_bytes) {
  if (slab_bytes < (1u << 20)) return fail(p->g, ZGPU_ERR_ARG, "a pass holds at least 1 MiB of slabs");
  p->pass_slab = slab_bytes;
  p->tables_ready = false;  // (the passes are laid out again at the next call)
  plan_drop_graphs(p);
  return ZGPU_OK;
}

int zgpu_plan_profile(zgpu_plan_t* p, int on) {
  p->profile = on != 0;
  return ZGPU_OK;
}

int zgpu_plan_stage_ms(zgpu_plan_t* p, float* ms, int cap) {
  if (p->profiled_op < 0) return fail(p->g, ZGPU_ERR_ARG, "no profiled call (zgpu_plan_profile)");
  const uint32_t first = p->profiled_op == 0 ? 0u : (uint32_t)P_DSTREAMS, nst = p->profiled_op == 0 ? 10u : 2u;
  for (int k = 0; k < cap; ++k) ms[k] = 0;
  for (size_t ps = 0; ps < p->passes.size(); ++ps)
    for (uint32_t s = 0; s < nst; ++s) {
      float t = 0;
      const size_t e = ps * nst + s;
      HIPCHK(p->g, hipEventSynchronize(p->stage_ev[e + 1]));
      HIPCHK(p->g, hipEventElapsedTime(&t, p->stage_ev[e], p->stage_ev[e + 1]));
      if ((int)(first + s) < cap) ms[first + s] += t;
    }
  return (int)P_STAGES;
}

int zgpu_plan_passes(const zgpu_plan_t* p) {
  if (!p->tables_ready) return -1;
  return (int)p->passes.size();
}

int zgpu_plan_set_stored(zgpu_plan_t* p, const uint64_t* comp_off, const uint32_t* comp_len, const uint8_t* method) {
  zgpu_t* g = p->g;
  if (!p->nb) return ZGPU_OK;
  std::vector<Block> tab(p->nb);
  for (uint32_t i = 0; i < p->nb; ++i) {
    const uint32_t b = p->order[i];
    tab[i] = Block{comp_off[b], p->htab[i].dst, p->htab[i].raw_len, comp_len[b], method[b], 0};
  }
  if (!p->d_dec) HIPCHK(g, hipMalloc((void**)&p->d_dec, p->nb * sizeof(Block)));
  HIPCHK(g, hipMemcpy(p->d_dec, tab.data(), p->nb * sizeof(Block), hipMemcpyHostToDevice));
  for (auto it = p->graphs.begin(); it != p->graphs.end();) {  // (decode graphs read the old table: rebuild them)
    if (std::get<0>(it->first) == 1) { (void)hipGraphExecDestroy(it->second); it = p->graphs.erase(it); }
    else ++it;
  }
  return ZGPU_OK;
}

int zgpu_plan_use_graphs(zgpu_plan_t* p, int on) {
  p->use_graphs = on != 0;
  return ZGPU_OK;
}

int zgpu_plan_encode(zgpu_plan_t* p, const uint8_t* d_in, int level, uint8_t* d_slabs, uint32_t* d_out_len) {
  if (level < 1 || level > 9) return fail(p->g, ZGPU_ERR_ARG, "level 1..9");
  return plan_call(p, 0, d_in, d_slabs, d_out_len, level);
}

int zgpu_plan_decode(zgpu_plan_t* p, const uint8_t* d_comp, uint8_t* d_out, uint32_t* d_err) {
  if (!p->d_dec) return fail(p->g, ZGPU_ERR_ARG, "zgpu_plan_set_stored first");
  return plan_call(p, 1, d_comp, d_out, d_err, 0);
}

// Two searches, each timed on the caller's data:
//   - per class, its bucket bits (hbits - 2 .. hbits + 1) and parse walkers (128 .. 1024), on the plain path's
//     launches over that class's blocks (a class's shape decides only the work, never the bytes);
//   - per stage, the persistent grid's workgroups per compute unit (1 .. 32, as resources allow), from whole
//     encodes of the plan with each setting, every stage's time taken from events between the stages.
int zgpu_plan_tune(zgpu_plan_t* p, const uint8_t* d_in, int level, uint8_t* d_slabs, uint32_t* d_out_len) {
  zgpu_t* g = p->g;
  if (level < 1 || level > 9) return fail(g, ZGPU_ERR_ARG, "level 1..9");
  if (!p->nb) return ZGPU_OK;
  hipEvent_t e0, e1;
  HIPCHK(g, hipEventCreate(&e0));
  HIPCHK(g, hipEventCreate(&e1));
  int rc = ZGPU_OK;
  DBuf work;
  for (size_t ci = 0; ci < p->cls.size() && rc == ZGPU_OK; ++ci) {
    PlanClass& c = p->cls[ci];
    const uint32_t count = std::min<uint32_t>(c.count, std::max<uint32_t>(1, (uint32_t)((1ull << 29) >> c.shape.blog)));
    const EncShape base = class_shape(c.shape.blog, p->key_log);
    EncShape best = c.shape;
    float best_ms = 1e30f;
    for (int dh = -2; dh <= 1 && rc == ZGPU_OK; ++dh) {
      for (uint32_t w : {128u, 256u, 512u, 1024u}) {
        EncShape S = base;
        S.hbits = (uint32_t)((int)base.hbits + dh);
        S.walkers = w;
        if (!shape_fits(S)) continue;
        if ((rc = ensure(g, work, enc_layout(count, S.blog, S.hbits).total))) break;
        float ms = 0;
        for (int rep = 0; rep < 2 && rc == ZGPU_OK; ++rep) {  // (the first run warms caches and code)
          (void)hipEventRecord(e0, g->stream);
          rc = encode_run(g, d_in, p->d_enc + c.first, count, S, level, d_slabs + p->slab_off[c.first],
                          p->d_tmp + c.first, (uint8_t*)work.p, g->stream, false);
          (void)hipEventRecord(e1, g->stream);
          if (rc == ZGPU_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(g, ZGPU_ERR_DEVICE, "tune");
          if (rc == ZGPU_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (rc == ZGPU_OK && ms < best_ms) { best_ms = ms; best = S; }
      }
    }
    c.shape = best;
  }
  if (work.p) (void)hipFree(work.p);
  plan_drop_graphs(p);
  if (rc == ZGPU_OK) rc = plan_tables(p);  // (the shapes changed: arenas and tables again)
  // the stages' grids
  const size_t nev = 1 + p->passes.size() * 10;
  while (rc == ZGPU_OK && p->stage_ev.size() < nev) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { rc = fail(g, ZGPU_ERR_DEVICE, "tune: events"); break; }
    p->stage_ev.push_back(e);
  }
  static const uint32_t OCC[] = {1, 2, 4, 8, 16, 32};
  float best_ms[P_STAGES];
  uint32_t best_occ[P_STAGES];
  for (uint32_t s = 0; s < P_STAGES; ++s) { best_ms[s] = 1e30f; best_occ[s] = p->occ[s]; }
  for (uint32_t oi = 0; oi < sizeof(OCC) / sizeof(OCC[0]) && rc == ZGPU_OK; ++oi) {
    for (uint32_t s = 0; s < P_STAGES; ++s) p->occ[s] = std::max<uint32_t>(1, std::min(OCC[oi], 32u));
    for (int rep = 0; rep < 2 && rc == ZGPU_OK; ++rep) {
      (void)hipEventRecord(p->stage_ev[0], g->stream);
      rc = plan_launch(p, 0, d_in, d_slabs, d_out_len, level, p->stage_ev.data() + 1);
      if (rc == ZGPU_OK && hipEventSynchronize(p->stage_ev[nev - 1]) != hipSuccess)
        rc = fail(g, ZGPU_ERR_DEVICE, "tune: a stage failed (its grid may not fit the GPU)");
      if (rc != ZGPU_OK || rep == 0) continue;
      float per[P_STAGES] = {};
      for (size_t ps = 0; ps < p->passes.size(); ++ps)
        for (uint32_t s = 0; s < 10; ++s) {
          float ms = 0;
          (void)hipEventElapsedTime(&ms, p->stage_ev[ps * 10 + s], p->stage_ev[ps * 10 + s + 1]);
          per[s] += ms;
        }
      for (uint32_t s = 0; s < 10; ++s)
        if (per[s] < best_ms[s]) { best_ms[s] = per[s]; best_occ[s] = p->occ[s]; }
    }
  }
  for (uint32_t s = 0; s < 10; ++s) p->occ[s] = best_occ[s];
  for (uint32_t s = 10; s < P_STAGES; ++s) p->occ[s] = DEFAULT_OCC[s];
  // the match stage's workgroup size, at the number of wavefronts per CU found best for 256 threads
  {
    const uint32_t occ256 = p->match_wg == 256 ? p->occ[P_MATCH] : p->occ[P_MATCH] * p->match_wg / 256;
    uint32_t best_w = p->match_wg, best_o = p->occ[P_MATCH];
    float best = 1e30f;
    for (uint32_t w : {64u, 128u, 256u, 512u, 1024u}) {
      if (rc != ZGPU_OK) break;
      p->match_wg = w;
      p->occ[P_MATCH] = std::max<uint32_t>(1, std::min<uint32_t>(32, occ256 * 256 / w));
      for (int rep = 0; rep < 2 && rc == ZGPU_OK; ++rep) {
        (void)hipEventRecord(p->stage_ev[0], g->stream);
        rc = plan_launch(p, 0, d_in, d_slabs, d_out_len, level, p->stage_ev.data() + 1);
        if (rc == ZGPU_OK && hipEventSynchronize(p->stage_ev[nev - 1]) != hipSuccess)
          rc = fail(g, ZGPU_ERR_DEVICE, "tune: the match stage failed");
        if (rc != ZGPU_OK || rep == 0) continue;
        float ms = 0;
        for (size_t ps = 0; ps < p->passes.size(); ++ps) {
          float m = 0;
          (void)hipEventElapsedTime(&m, p->stage_ev[ps * 10 + P_MATCH], p->stage_ev[ps * 10 + P_MATCH + 1]);
          ms += m;
        }
        if (ms < best) { best = ms; best_w = w; best_o = p->occ[P_MATCH]; }
      }
    }
    p->match_wg = best_w;
    p->occ[P_MATCH] = best_o;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  plan_drop_graphs(p);
  return rc;
}

}  // extern "C"
