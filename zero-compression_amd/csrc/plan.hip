// The batched form: the blocks of many files, of uneven lengths, planned together (included by capi.hip).
//
// A block of r bytes runs in a slab of 2^c bytes, c = max(10, ceil(log2 r)) up to the archive's block_log (its size
// class): the position-parallel kernels then launch work for 2^c positions instead of 2^block_log, and the per-block
// scratch shrinks the same way. The hash keys keep the archive's width, so every block's payload is the one the
// plain path writes. Each class is cut into jobs of at most 2^29 slab bytes, and the jobs are spread over a few
// streams (longest first), each stream with its own scratch, so the small classes overlap the large ones.
// Optionally the whole launch sequence of a call is captured into a hipGraph the first time and replayed when the
// same buffers come again. zgpu_plan_tune times the output-invariant shape parameters of each class (bucket bits,
// parse walkers) on the caller's data and keeps the fastest.
#include <algorithm>
#include <map>
#include <tuple>

namespace {

__global__ void k_scatter_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                              const uint32_t* __restrict__ order, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[order[i]] = src[i];
}

struct PlanClass {
  EncShape shape;
  uint32_t first, count;  // plan-order block range
};
struct PlanJob {
  uint32_t cls, first, count;
  int stream;
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
size_t parse_lds_bytes(const EncShape& S) {
  const size_t B = (size_t)1 << S.blog;
  return (B / 32 + 19 * (size_t)S.walkers) * 4;
}
bool shape_fits(const EncShape& S) {
  const size_t np = (size_t)1 << pbits_of(S.blog, S.hbits);
  return shape_ok(S) && parse_lds_bytes(S) <= 160 * 1024 && (((size_t)1 << S.hbits) + 16 + 17 * np) * 4 <= 160 * 1024 &&
         (((size_t)1 << (S.hbits - pbits_of(S.blog, S.hbits))) + PLACE_CAP) * 4 <= 160 * 1024;
}

}  // namespace

struct zgpu_plan {
  zgpu_t* g = nullptr;
  uint32_t nb = 0, key_log = 18;
  std::vector<uint32_t> order;     // plan index -> caller's block index
  std::vector<uint64_t> slab_off;  // plan index -> slab offset
  uint64_t slab_total = 0;
  std::vector<PlanClass> cls;
  std::vector<PlanJob> jobs;
  std::vector<hipStream_t> streams;
  std::vector<DBuf> ework, dwork;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  std::vector<Block> htab;  // plan order: src, dst = src, raw_len
  Block* d_enc = nullptr;  // plan order
  Block* d_dec = nullptr;  // plan order (after zgpu_plan_set_stored)
  uint32_t* d_order = nullptr;
  uint32_t* d_tmp = nullptr;
  bool use_graphs = false;
  std::map<std::tuple<int, const void*, const void*, const void*, int>, hipGraphExec_t> graphs;
};

static void plan_free(zgpu_plan* p) {
  for (auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  p->graphs.clear();
  for (DBuf& b : p->ework) if (b.p) (void)hipFree(b.p);
  for (DBuf& b : p->dwork) if (b.p) (void)hipFree(b.p);
  for (hipStream_t s : p->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : p->join) (void)hipEventDestroy(e);
  if (p->fork) (void)hipEventDestroy(p->fork);
  for (void* q : {(void*)p->d_enc, (void*)p->d_dec, (void*)p->d_order, (void*)p->d_tmp})
    if (q) (void)hipFree(q);
  delete p;
}

// spreads the jobs over the streams, largest first, each to the least loaded stream
static void plan_assign(zgpu_plan* p) {
  std::vector<size_t> idx(p->jobs.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  auto cost = [&](size_t i) { return (uint64_t)p->jobs[i].count << p->cls[p->jobs[i].cls].shape.blog; };
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
  std::vector<uint64_t> load(p->streams.size(), 0);
  for (size_t i : idx) {
    const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    p->jobs[i].stream = s;
    load[s] += cost(i);
  }
}

static int plan_workspace(zgpu_plan* p) {
  std::vector<size_t> en(p->streams.size(), 0), de(p->streams.size(), 0);
  for (const PlanJob& j : p->jobs) {
    const EncShape& S = p->cls[j.cls].shape;
    en[j.stream] = std::max(en[j.stream], enc_layout(j.count, S.blog, S.hbits).total);
    de[j.stream] = std::max(de[j.stream], dec_layout(j.count, S.blog).total);
  }
  for (size_t s = 0; s < p->streams.size(); ++s) {
    if (int rc = ensure(p->g, p->ework[s], en[s])) return rc;
    if (int rc = ensure(p->g, p->dwork[s], de[s])) return rc;
  }
  return ZGPU_OK;
}

// the launch sequence of one call: fork to the streams, every job, join, put out_len / err in the caller's order
static int plan_launch(zgpu_plan* p, int op, const uint8_t* d_src, uint8_t* d_dst, uint32_t* d_res, int level) {
  zgpu_t* g = p->g;
  hipStream_t main = g->stream;
  HIPCHK(g, hipEventRecord(p->fork, main));
  for (hipStream_t s : p->streams) HIPCHK(g, hipStreamWaitEvent(s, p->fork, 0));
  for (const PlanJob& j : p->jobs) {
    const PlanClass& c = p->cls[j.cls];
    hipStream_t st = p->streams[j.stream];
    int rc;
    if (op == 0)
      rc = encode_run(g, d_src, p->d_enc + j.first, j.count, c.shape, level, d_dst + p->slab_off[j.first],
                      p->d_tmp + j.first, (uint8_t*)p->ework[j.stream].p, st, false);
    else
      rc = decode_run(g, d_src, p->d_dec + j.first, j.count, c.shape.blog, d_dst, p->d_tmp + j.first,
                      (uint8_t*)p->dwork[j.stream].p, st, false);
    if (rc) return rc;
  }
  for (size_t s = 0; s < p->streams.size(); ++s) {
    HIPCHK(g, hipEventRecord(p->join[s], p->streams[s]));
    HIPCHK(g, hipStreamWaitEvent(main, p->join[s], 0));
  }
  k_scatter_u32<<<(p->nb + 255) / 256, 256, 0, main>>>(d_res, p->d_tmp, p->d_order, p->nb);
  HIPCHK(g, hipGetLastError());
  return ZGPU_OK;
}

static int plan_call(zgpu_plan* p, int op, const uint8_t* d_src, uint8_t* d_dst, uint32_t* d_res, int level) {
  if (p->nb == 0) return ZGPU_OK;
  if (int rc = plan_workspace(p)) return rc;
  if (!p->use_graphs) return plan_launch(p, op, d_src, d_dst, d_res, level);
  const auto key = std::make_tuple(op, (const void*)d_src, (const void*)d_dst, (const void*)d_res, level);
  auto it = p->graphs.find(key);
  if (it == p->graphs.end()) {
    zgpu_t* g = p->g;
    hipGraph_t graph;
    HIPCHK(g, hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    const int rc = plan_launch(p, op, d_src, d_dst, d_res, level);
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
  zgpu_plan* p = new zgpu_plan();
  p->g = g;
  p->nb = nb;
  p->key_log = key_log;
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
    PlanClass pc{class_shape(c, key_log), i, j - i};
    p->cls.push_back(pc);
    const uint32_t per_job = std::max<uint32_t>(1, (uint32_t)(((uint64_t)1 << 29) >> c));
    for (uint32_t k = i; k < j; k += per_job)
      p->jobs.push_back({(uint32_t)p->cls.size() - 1, k, std::min(per_job, j - k), 0});
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
  p->streams.resize(streams);
  p->join.resize(streams);
  p->ework.resize(streams);
  p->dwork.resize(streams);
  for (int s = 0; s < streams; ++s) {
    chk(hipStreamCreateWithFlags(&p->streams[s], hipStreamNonBlocking), "hipStreamCreateWithFlags");
    chk(hipEventCreateWithFlags(&p->join[s], hipEventDisableTiming), "hipEventCreate");
  }
  chk(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming), "hipEventCreate");
  plan_assign(p);
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

int zgpu_plan_tune(zgpu_plan_t* p, const uint8_t* d_in, int level, uint8_t* d_slabs, uint32_t* d_out_len) {
  zgpu_t* g = p->g;
  if (level < 1 || level > 9) return fail(g, ZGPU_ERR_ARG, "level 1..9");
  hipEvent_t e0, e1;
  HIPCHK(g, hipEventCreate(&e0));
  HIPCHK(g, hipEventCreate(&e1));
  int rc = ZGPU_OK;
  for (size_t ci = 0; ci < p->cls.size() && rc == ZGPU_OK; ++ci) {
    PlanClass& c = p->cls[ci];
    const PlanJob* job = nullptr;
    for (const PlanJob& j : p->jobs) if (j.cls == ci) { job = &j; break; }
    if (!job) continue;
    const EncShape base = class_shape(c.shape.blog, p->key_log);
    EncShape best = c.shape;
    float best_ms = 1e30f;
    for (int dh = -2; dh <= 1 && rc == ZGPU_OK; ++dh) {
      for (uint32_t w : {128u, 256u, 512u, 1024u}) {
        EncShape S = base;
        S.hbits = (uint32_t)((int)base.hbits + dh);
        S.walkers = w;
        if (!shape_fits(S)) continue;
        DBuf& work = p->ework[0];
        if ((rc = ensure(g, work, enc_layout(job->count, S.blog, S.hbits).total))) break;
        float ms = 0;
        for (int rep = 0; rep < 2 && rc == ZGPU_OK; ++rep) {  // (the first run warms caches and code)
          (void)hipEventRecord(e0, g->stream);
          rc = encode_run(g, d_in, p->d_enc + job->first, job->count, S, level, d_slabs + p->slab_off[job->first],
                          p->d_tmp + job->first, (uint8_t*)work.p, g->stream, false);
          (void)hipEventRecord(e1, g->stream);
          if (rc == ZGPU_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(g, ZGPU_ERR_DEVICE, "tune");
          if (rc == ZGPU_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (rc == ZGPU_OK && ms < best_ms) { best_ms = ms; best = S; }
      }
    }
    c.shape = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);  // (the shapes changed)
  p->graphs.clear();
  if (rc) return rc;
  (void)d_out_len;
  return ZGPU_OK;
}

}  // extern "C"
