// The serving loop (included by capi.hip): compress and decompress requests of mixed sizes and levels, admitted by
// HBM budget into batches that run while later requests keep arriving (continuous batching). The reference has no
// server; this is the form a service wraps around the codec.
//
// A Server holds `slots` batch slots, each with its own stream, device arenas and encoder/decoder scratch, sized
// once from the HBM budget, so nothing is allocated while serving. When a slot is free, every queued request that
// fits the slot's arenas joins one batch, first come first served. A batch runs in two phases on its slot's stream:
//   1. inputs and block tables to the device; the compress blocks by size class (a block of r bytes in a slab of
//      2^c, c = 10, 13, 16 or block_log), each block at its own level; the decompress blocks; the decompressed bytes
//      and the payload lengths back to the host. Decompress requests are done here.
//   2. the payloads of each compress request gathered back to back on the device (the host has the lengths now) and
//      copied into the request's output. Compress requests are done here.
// Slots overlap each other: one batch's copies run beside another's kernels. Every result depends only on its own
// request: the encoder is deterministic, so a request gives the same bytes alone and in any mix.
#include <chrono>
#include <cmath>
#include <deque>
#include <random>

namespace {

// payload of encoded block b (at slab + soff[b], len[b] bytes, or none when stored) to pay + goff[b]
__global__ __launch_bounds__(256) void k_gather(const uint8_t* __restrict__ slab, const uint64_t* __restrict__ soff,
                                                const uint32_t* __restrict__ len, const uint64_t* __restrict__ goff,
                                                uint8_t* __restrict__ pay) {
  const uint32_t b = blockIdx.x;
  const uint32_t n = len[b];
  if (n == 0xFFFFFFFFu) return;
  const uint8_t* s = slab + soff[b];
  uint8_t* d = pay + goff[b];
  const uint32_t full = n & ~15u;
  for (uint32_t i = threadIdx.x * 16; i < full; i += 256 * 16) {
    uint4 v;
    __builtin_memcpy(&v, s + i, 16);
    __builtin_memcpy(d + i, &v, 16);
  }
  if (full + threadIdx.x < n) d[full + threadIdx.x] = s[full + threadIdx.x];
}

// one piece of a copy list: len bytes from src to dst. Either side may be pinned host memory, which the device reaches
// over PCIe at the same address as the host (checked when it is registered with the server).
struct CopyPiece {
  uint64_t src, dst;
  uint32_t len, pad;
};
constexpr uint32_t PIECE = 1u << 18;  // the largest piece: a request's bytes are cut into pieces of at most this
// One workgroup per piece: 16 bytes per thread, 4 loads in flight per thread, then the tail (under 16 bytes) byte by
// byte. A batch's requests thus move in one launch per direction instead of one copy call each.
__global__ __launch_bounds__(256) void k_copy_list(const CopyPiece* __restrict__ pcs) {
  const CopyPiece pc = pcs[blockIdx.x];
  const uint8_t* s = (const uint8_t*)pc.src;
  uint8_t* d = (uint8_t*)pc.dst;
  const uint32_t n = pc.len, full = n & ~15u;
  uint32_t i = threadIdx.x * 16;
  for (; i + 3 * 256 * 16 < full; i += 4 * 256 * 16) {
    uint4 v[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) __builtin_memcpy(&v[k], s + i + k * 256 * 16, 16);
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) __builtin_memcpy(d + i + k * 256 * 16, &v[k], 16);
  }
  for (; i < full; i += 256 * 16) {
    uint4 v;
    __builtin_memcpy(&v, s + i, 16);
    __builtin_memcpy(d + i, &v, 16);
  }
  if (full + threadIdx.x < n) d[full + threadIdx.x] = s[full + threadIdx.x];
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SReq {
  int op = 0;     // 0 compress, 1 decompress
  int level = 5;  // compress
  const uint8_t* in = nullptr;  // compress: raw bytes; decompress: the stored blocks back to back (pinned host)
  uint64_t in_len = 0;
  uint64_t raw_len = 0;
  uint32_t nblk = 0;
  const uint32_t* clen = nullptr;  // decompress: per block
  const uint8_t* meth = nullptr;
  uint8_t* out = nullptr;      // compress: payloads of its compressed blocks back to back; decompress: raw bytes
  uint32_t* out_len = nullptr; // compress: per block, payload size or ZGPU_STORED
  uint64_t out_bytes = 0;      // compress: payload bytes written
  double t_arr = 0, t_done = 0;
  uint32_t err = 0;
  // placement in the batch
  uint64_t ioff = 0, roff = 0;
  uint32_t first_blk = 0;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  int phase = 0;  // 0 free, 1 kernels and copies in flight, 2 gather in flight
  std::vector<SReq*> reqs;
  uint8_t *d_in = nullptr, *d_slab = nullptr, *d_pay = nullptr, *d_raw = nullptr;
  Block* d_tab = nullptr;
  uint32_t *d_len = nullptr, *d_err = nullptr;
  uint64_t *d_soff = nullptr, *d_goff = nullptr;
  DBuf ework, dwork;
  Block* h_tab = nullptr;
  uint32_t *h_len = nullptr, *h_err = nullptr;
  uint64_t *h_soff = nullptr, *h_goff = nullptr;
  // copy lists: A holds a batch's inputs (phase 1) and later its compress payloads (phase 2), B its decompressed
  // bytes (phase 1); pinned host tables and their device copies
  CopyPiece *h_pa = nullptr, *h_pb = nullptr, *d_pa = nullptr, *d_pb = nullptr;
  uint32_t nenc = 0, ndec = 0;
};

struct Server {
  zgpu_t* g = nullptr;
  uint32_t blog = 18;
  std::vector<uint32_t> cls;  // slab logs of the size classes: 10, 13, 16 (those below block_log), block_log
  uint64_t cap_in = 0, cap_slab = 0, cap_raw = 0, hbm = 0;
  uint32_t cap_blk = 0, cap_dec = 0, job_blocks_log = 0, cap_pieces = 0;
  std::vector<std::pair<uint64_t, uint64_t>> host;  // pinned host ranges [a, e) that requests may read or write
  std::vector<Slot> slots;
  std::deque<SReq*> queue;
  uint64_t batches = 0, batched_reqs = 0, completed = 0;

  int init(zgpu_t* gg, uint64_t budget, int nslots, uint32_t block_log) {
    g = gg;
    blog = block_log;
    for (uint32_t c : {10u, 13u, 16u})
      if (c < blog) cls.push_back(c);
    cls.push_back(blog);
    const uint64_t per = budget / (uint64_t)nslots;
    // per slot: inputs I, slabs 2I, gathered payloads I, decompressed bytes I, the decoder's scratch (about 2I),
    // and the encoder's scratch for jobs of 2^28 slab positions (about 19 bytes each)
    job_blocks_log = 28;  // (fewer positions per encoder job when the budget is small: at most a quarter of it)
    while (job_blocks_log > blog &&
           enc_layout((uint32_t)(((uint64_t)1 << job_blocks_log) >> blog), blog, hbits_of(blog)).total > per / 4)
      --job_blocks_log;
    const uint64_t enc_ws = enc_layout((uint32_t)(((uint64_t)1 << job_blocks_log) >> blog), blog, hbits_of(blog)).total;
    const uint64_t spare = per > enc_ws + ((uint64_t)1 << 29) ? per - enc_ws - ((uint64_t)1 << 29) : 0;
    const uint64_t I = std::max<uint64_t>((uint64_t)64 << 20, spare / 9);
    cap_in = I;
    cap_slab = 2 * I + ((uint64_t)1 << 20);
    cap_raw = I;
    cap_blk = (uint32_t)std::min<uint64_t>(1u << 20, I / 1024 + (I >> blog) + 16);
    cap_dec = (uint32_t)std::min<uint64_t>(cap_blk, (I >> blog) + 1024);
    cap_pieces = cap_blk + (uint32_t)(I / PIECE) + 16;  // (every request has a block; at most I bytes move per list)
    slots.resize(nslots);
    for (Slot& s : slots) {
      HIPCHK(g, hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
      HIPCHK(g, hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
      HIPCHK(g, hipMalloc((void**)&s.d_in, cap_in));
      HIPCHK(g, hipMalloc((void**)&s.d_slab, cap_slab));
      HIPCHK(g, hipMalloc((void**)&s.d_pay, cap_in));
      HIPCHK(g, hipMalloc((void**)&s.d_raw, cap_raw));
      HIPCHK(g, hipMalloc((void**)&s.d_tab, (size_t)cap_blk * sizeof(Block)));
      HIPCHK(g, hipMalloc((void**)&s.d_len, (size_t)cap_blk * 4));
      HIPCHK(g, hipMalloc((void**)&s.d_err, (size_t)cap_blk * 4));
      HIPCHK(g, hipMalloc((void**)&s.d_soff, (size_t)cap_blk * 8));
      HIPCHK(g, hipMalloc((void**)&s.d_goff, (size_t)cap_blk * 8));
      HIPCHK(g, hipHostMalloc((void**)&s.h_tab, (size_t)cap_blk * sizeof(Block), hipHostMallocDefault));
      HIPCHK(g, hipHostMalloc((void**)&s.h_len, (size_t)cap_blk * 4, hipHostMallocDefault));
      HIPCHK(g, hipHostMalloc((void**)&s.h_err, (size_t)cap_blk * 4, hipHostMallocDefault));
      HIPCHK(g, hipHostMalloc((void**)&s.h_soff, (size_t)cap_blk * 8, hipHostMallocDefault));
      HIPCHK(g, hipHostMalloc((void**)&s.h_goff, (size_t)cap_blk * 8, hipHostMallocDefault));
      HIPCHK(g, hipHostMalloc((void**)&s.h_pa, (size_t)cap_pieces * sizeof(CopyPiece), hipHostMallocDefault));
      HIPCHK(g, hipHostMalloc((void**)&s.h_pb, (size_t)cap_pieces * sizeof(CopyPiece), hipHostMallocDefault));
      HIPCHK(g, hipMalloc((void**)&s.d_pa, (size_t)cap_pieces * sizeof(CopyPiece)));
      HIPCHK(g, hipMalloc((void**)&s.d_pb, (size_t)cap_pieces * sizeof(CopyPiece)));
      // the largest scratch any job of this slot can need, allocated now
      size_t ew = 0;
      for (uint32_t cl : cls) {
        const EncShape S = class_shape(cl, blog);
        ew = std::max(ew, enc_layout((uint32_t)(((uint64_t)1 << job_blocks_log) >> cl), cl, S.hbits).total);
      }
      if (int rc = ensure(g, s.ework, ew)) return rc;
      if (int rc = ensure(g, s.dwork, dec_layout(cap_dec, blog).total)) return rc;
      hbm += 2 * cap_in + cap_slab + cap_raw + (size_t)cap_blk * (sizeof(Block) + 24) + ew + s.dwork.n +
             2 * (size_t)cap_pieces * sizeof(CopyPiece);
    }
    return ZGPU_OK;
  }

  void release() {
    for (Slot& s : slots) {
      (void)hipStreamSynchronize(s.st);
      for (void* p : {(void*)s.d_in, (void*)s.d_slab, (void*)s.d_pay, (void*)s.d_raw, (void*)s.d_tab,
                      (void*)s.d_len, (void*)s.d_err, (void*)s.d_soff, (void*)s.d_goff, (void*)s.d_pa,
                      (void*)s.d_pb, s.ework.p, s.dwork.p})
        if (p) (void)hipFree(p);
      for (void* p : {(void*)s.h_tab, (void*)s.h_len, (void*)s.h_err, (void*)s.h_soff, (void*)s.h_goff,
                      (void*)s.h_pa, (void*)s.h_pb})
        if (p) (void)hipHostFree(p);
      if (s.st) (void)hipStreamDestroy(s.st);
      if (s.ev) (void)hipEventDestroy(s.ev);
    }
    slots.clear();
  }

  // registers pinned host memory that requests read from or write to; the device must reach it at the host's address
  int add_host(const void* p, uint64_t n) {
    hipPointerAttribute_t a;
    HIPCHK(g, hipPointerGetAttributes(&a, p));
    if (a.type != hipMemoryTypeHost || a.devicePointer != p)
      return fail(g, ZGPU_ERR_ARG, "serve: request buffers must be pinned host memory mapped at the same address");
    host.push_back({(uint64_t)(uintptr_t)p, (uint64_t)(uintptr_t)p + n});
    return ZGPU_OK;
  }
  bool host_ok(const void* p, uint64_t n) const {
    const uint64_t a = (uint64_t)(uintptr_t)p;
    for (const auto& h : host)
      if (a >= h.first && a <= h.second && n <= h.second - a) return true;
    return n == 0;
  }
  // appends the pieces of a copy of n bytes from src to dst
  static uint32_t pieces(CopyPiece* pc, uint32_t np, const void* src, void* dst, uint64_t n) {
    for (uint64_t o = 0; o < n; o += PIECE)
      pc[np++] = CopyPiece{(uint64_t)(uintptr_t)src + o, (uint64_t)(uintptr_t)dst + o,
                           (uint32_t)std::min<uint64_t>(PIECE, n - o), 0u};
    return np;
  }
  // the copy list pc[0, np) as one launch on the slot's stream (its table goes up first)
  int run_list(Slot& s, CopyPiece* h, CopyPiece* d, uint32_t np) {
    if (!np) return ZGPU_OK;
    HIPCHK(g, hipMemcpyAsync(d, h, (size_t)np * sizeof(CopyPiece), hipMemcpyHostToDevice, s.st));
    k_copy_list<<<np, 256, 0, s.st>>>(d);
    HIPCHK(g, hipGetLastError());
    return ZGPU_OK;
  }

  // the size class of a block of raw bytes: the index of the smallest class slab that holds it
  uint32_t class_of(uint32_t raw) const {
    uint32_t k = 0;
    while (k + 1 < cls.size() && raw > (1u << cls[k])) ++k;
    return k;
  }
  uint32_t nblocks(const SReq* r) const {
    return r->op == 0 ? (uint32_t)((r->raw_len + (1ull << blog) - 1) >> blog) : r->nblk;
  }
  uint64_t slab_need(const SReq* r) const {
    uint64_t t = 0;
    const uint64_t B = 1ull << blog;
    for (uint64_t o = 0; o < r->raw_len; o += B) t += 1ull << cls[class_of((uint32_t)std::min(B, r->raw_len - o))];
    return t;
  }

  // forms a batch in slot s from the head of the queue and starts phase 1
  int launch(Slot& s) {
    uint64_t in = 0, slab = 0, raw = 0;
    uint32_t nenc = 0, ndec = 0;
    s.reqs.clear();
    while (!queue.empty()) {
      SReq* r = queue.front();
      const uint32_t nbk = nblocks(r);
      const uint64_t sl = r->op == 0 ? slab_need(r) : 0;
      const uint64_t rw = r->op == 1 ? r->raw_len : 0;
      const uint64_t inl = (r->in_len + 255) & ~(uint64_t)255;
      const bool bad_buf = !host_ok(r->in, r->in_len) || !host_ok(r->out, r->raw_len);
      if (bad_buf || in + inl > cap_in || slab + sl > cap_slab || raw + rw > cap_raw ||
          nenc + ndec + nbk > cap_blk || (r->op == 1 && ndec + nbk > cap_dec)) {
        if (s.reqs.empty() || bad_buf) {  // (it can never fit, or its buffers are not the server's: refused)
          r->err = ZGPU_E_METHOD;
          r->t_done = now_s();
          ++completed;
          queue.pop_front();
          continue;
        }
        break;
      }
      queue.pop_front();
      r->ioff = in;
      r->roff = raw;
      in += inl;
      slab += sl;
      raw += (rw + 255) & ~(uint64_t)255;
      if (r->op == 0) nenc += nbk; else ndec += nbk;
      s.reqs.push_back(r);
    }
    if (s.reqs.empty()) return ZGPU_OK;
    ++batches;
    batched_reqs += s.reqs.size();
    // block tables: the compress blocks grouped by class (rows [0, nenc)), then the decompress blocks
    const uint64_t B = 1ull << blog;
    const size_t K = cls.size();
    std::vector<uint32_t> count(K, 0), at(K, 0);
    std::vector<uint64_t> cbase(K, 0);
    for (SReq* r : s.reqs)
      if (r->op == 0)
        for (uint64_t o = 0; o < r->raw_len; o += B) ++count[class_of((uint32_t)std::min(B, r->raw_len - o))];
    for (size_t k = 1; k < K; ++k) {
      at[k] = at[k - 1] + count[k - 1];
      cbase[k] = cbase[k - 1] + ((uint64_t)count[k - 1] << cls[k - 1]);
    }
    std::vector<uint32_t> fill = at;
    uint32_t di = nenc, npa = 0;
    for (SReq* r : s.reqs) {
      npa = pieces(s.h_pa, npa, r->in, s.d_in + r->ioff, r->in_len);
      if (r->op == 0) {
        for (uint64_t o = 0; o < r->raw_len; o += B) {
          const uint32_t n = (uint32_t)std::min(B, r->raw_len - o);
          const uint32_t ci = class_of(n);
          const uint32_t i = fill[ci]++;
          s.h_tab[i] = Block{r->ioff + o, 0, n, 0, 0, (uint32_t)r->level};
          s.h_soff[i] = cbase[ci] + ((uint64_t)(i - at[ci]) << cls[ci]);
        }
      } else {
        uint64_t co = 0, ro = 0;
        r->first_blk = di;
        for (uint32_t k = 0; k < r->nblk; ++k) {
          const uint32_t n = (uint32_t)std::min(B, r->raw_len - ro);
          s.h_tab[di++] = Block{r->ioff + co, r->roff + ro, n, r->clen[k], r->meth[k], 0};
          co += r->clen[k];
          ro += n;
        }
      }
    }
    s.nenc = nenc;
    s.ndec = ndec;
    const uint32_t nt = nenc + ndec;
    if (int rc = run_list(s, s.h_pa, s.d_pa, npa)) return rc;  // the inputs, all in one launch
    HIPCHK(g, hipMemcpyAsync(s.d_tab, s.h_tab, (size_t)nt * sizeof(Block), hipMemcpyHostToDevice, s.st));
    if (nenc) HIPCHK(g, hipMemcpyAsync(s.d_soff, s.h_soff, (size_t)nenc * 8, hipMemcpyHostToDevice, s.st));
    // the encoder, class by class, in jobs of at most 2^28 slab positions
    for (size_t k = 0; k < K; ++k) {
      const uint32_t cl = cls[k];
      if (!count[k]) continue;
      const EncShape S = class_shape(cl, blog);
      const uint32_t per = std::max<uint32_t>(1, (uint32_t)(((uint64_t)1 << job_blocks_log) >> cl));
      for (uint32_t j = 0; j < count[k]; j += per) {
        const uint32_t n = std::min(per, count[k] - j);
        const uint32_t i0 = at[k] + j;
        if (int rc = encode_run(g, s.d_in, s.d_tab + i0, n, S, 5, s.d_slab + cbase[k] + ((uint64_t)j << cl),
                                s.d_len + i0, (uint8_t*)s.ework.p, s.st, false))
          return rc;
      }
    }
    if (ndec) {
      if (int rc = decode_run(g, s.d_in, s.d_tab + nenc, ndec, blog, s.d_raw, s.d_err + nenc, (uint8_t*)s.dwork.p,
                              s.st, false))
        return rc;
      HIPCHK(g, hipMemcpyAsync(s.h_err + nenc, s.d_err + nenc, (size_t)ndec * 4, hipMemcpyDeviceToHost, s.st));
      uint32_t npb = 0;
      for (SReq* r : s.reqs)
        if (r->op == 1) npb = pieces(s.h_pb, npb, s.d_raw + r->roff, r->out, r->raw_len);
      if (int rc = run_list(s, s.h_pb, s.d_pb, npb)) return rc;  // the decompressed bytes, in one launch
    }
    if (nenc) HIPCHK(g, hipMemcpyAsync(s.h_len, s.d_len, (size_t)nenc * 4, hipMemcpyDeviceToHost, s.st));
    HIPCHK(g, hipEventRecord(s.ev, s.st));
    s.phase = 1;
    return ZGPU_OK;
  }

  // phase 2: place each compress request's payloads back to back, gather them, copy them out
  int gather(Slot& s) {
    // the table row of each compress request's blocks, in block order (the fill of launch() again)
    const uint64_t B = 1ull << blog;
    const size_t K = cls.size();
    std::vector<uint32_t> row(s.nenc), count(K, 0), at(K, 0);
    for (SReq* r : s.reqs)
      if (r->op == 0)
        for (uint64_t o = 0; o < r->raw_len; o += B) ++count[class_of((uint32_t)std::min(B, r->raw_len - o))];
    for (size_t k = 1; k < K; ++k) at[k] = at[k - 1] + count[k - 1];
    uint32_t q = 0;
    for (SReq* r : s.reqs) {
      if (r->op != 0) continue;
      r->first_blk = q;
      for (uint64_t o = 0; o < r->raw_len; o += B) row[q++] = at[class_of((uint32_t)std::min(B, r->raw_len - o))]++;
    }
    for (SReq* r : s.reqs) {
      if (r->op != 0) continue;
      const uint32_t nbk = nblocks(r);
      uint64_t o = 0;
      for (uint32_t k = 0; k < nbk; ++k) {
        const uint32_t i = row[r->first_blk + k];
        const uint32_t l = s.h_len[i];
        r->out_len[k] = l;
        s.h_goff[i] = r->ioff + o;
        if (l != ZGPU_STORED) o += l;
      }
      r->out_bytes = o;
    }
    HIPCHK(g, hipMemcpyAsync(s.d_goff, s.h_goff, (size_t)s.nenc * 8, hipMemcpyHostToDevice, s.st));
    k_gather<<<s.nenc, 256, 0, s.st>>>(s.d_slab, s.d_soff, s.d_len, s.d_goff, s.d_pay);
    HIPCHK(g, hipGetLastError());
    uint32_t npa = 0;  // (list A is free again: phase 1, whose table it held, has completed)
    for (SReq* r : s.reqs)
      if (r->op == 0) npa = pieces(s.h_pa, npa, s.d_pay + r->ioff, r->out, r->out_bytes);
    if (int rc = run_list(s, s.h_pa, s.d_pa, npa)) return rc;  // the payloads, in one launch
    HIPCHK(g, hipEventRecord(s.ev, s.st));
    s.phase = 2;
    return ZGPU_OK;
  }

  // advances every slot whose work is done, and starts batches in the free slots
  int poll() {
    for (Slot& s : slots) {
      if (s.phase == 0) continue;
      const hipError_t q = hipEventQuery(s.ev);
      if (q == hipErrorNotReady) continue;
      HIPCHK(g, q);
      const double t = now_s();
      if (s.phase == 1) {
        for (SReq* r : s.reqs)
          if (r->op == 1) {
            uint32_t e = 0;
            for (uint32_t k = 0; k < r->nblk; ++k) e |= s.h_err[r->first_blk + k];
            r->err = e;
            r->t_done = t;
            ++completed;
          }
        bool any_enc = false;
        for (SReq* r : s.reqs) any_enc |= r->op == 0;
        if (any_enc) {
          if (int rc = gather(s)) return rc;
          continue;
        }
      } else {
        for (SReq* r : s.reqs)
          if (r->op == 0) { r->t_done = t; ++completed; }
      }
      s.phase = 0;
      s.reqs.clear();
    }
    for (Slot& s : slots)
      if (s.phase == 0 && !queue.empty())
        if (int rc = launch(s)) return rc;
    return ZGPU_OK;
  }
  bool idle() const {
    for (const Slot& s : slots) if (s.phase) return false;
    return queue.empty();
  }
};

}  // namespace

extern "C" {

int zgpu_serve_bench(zgpu_t* g, const zgpu_serve_cfg_t* cfg, const uint8_t* lz, uint64_t lz_len, const uint8_t* rnd,
                     uint64_t rnd_len, zgpu_serve_report_t* rep) {
  memset(rep, 0, sizeof *rep);
  const uint32_t blog = cfg->block_log;
  if (blog < 16 || blog > 20 || cfg->slots < 1 || cfg->slots > 8 || lz_len < (1ull << blog) ||
      rnd_len < (1ull << blog))
    return fail(g, ZGPU_ERR_ARG, "serve: block_log 16..20, slots 1..8, pools of at least one block");
  const uint64_t B = 1ull << blog;
  int rc = ZGPU_OK;
  // ---- the pools in pinned memory, and their blocks encoded once at level 5 (the decompress requests' inputs)
  struct Pool {
    const uint8_t* src;
    uint64_t n;
    uint8_t* raw = nullptr;   // pinned copy
    uint8_t* comp = nullptr;  // pinned: the stored blocks back to back
    std::vector<uint64_t> coff;
    std::vector<uint32_t> clen;
    std::vector<uint8_t> meth;
  } pools[2] = {{lz, lz_len}, {rnd, rnd_len}};
  for (int pi = 0; pi < 2; ++pi) {
    Pool& P = pools[pi];
    HIPCHK(g, hipHostMalloc((void**)&P.raw, P.n, hipHostMallocDefault));
    memcpy(P.raw, P.src, P.n);
    const uint32_t nb = (uint32_t)((P.n + B - 1) / B);
    P.coff.resize(nb + 1);
    P.clen.resize(nb);
    P.meth.resize(nb);
    if (cfg->dec_comp[pi]) {  // compressed by the caller
      if (!cfg->dec_comp_len[pi] || !cfg->dec_method[pi])
        return fail(g, ZGPU_ERR_ARG, "serve: dec_comp needs dec_comp_len and dec_method");
      uint64_t t = 0;
      for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t raw = (uint32_t)std::min<uint64_t>(B, P.n - (uint64_t)b * B);
        P.meth[b] = cfg->dec_method[pi][b];
        P.clen[b] = cfg->dec_comp_len[pi][b];
        if (P.meth[b] > 1 || (P.meth[b] == 0 && P.clen[b] != raw))
          return fail(g, ZGPU_ERR_ARG, "serve: a caller-compressed block has an impossible method or length");
        P.coff[b] = t;
        t += P.clen[b];
      }
      P.coff[nb] = t;
      HIPCHK(g, hipHostMalloc((void**)&P.comp, t ? t : 1, hipHostMallocDefault));
      memcpy(P.comp, cfg->dec_comp[pi], t);
      continue;
    }
    std::vector<uint32_t> len(nb);
    uint8_t* pays = (uint8_t*)malloc((size_t)P.n + 1);
    if (!pays) return fail(g, ZGPU_ERR_NOMEM, "serve: host memory");
    uint64_t used = 0;
    for (uint32_t b0 = 0; b0 < nb && rc == ZGPU_OK; b0 += 4096) {
      const uint32_t k = std::min<uint32_t>(4096, nb - b0);
      const uint64_t o = (uint64_t)b0 * B, n = std::min<uint64_t>((uint64_t)k * B, P.n - o);
      rc = zgpu_compress(g, P.raw + o, n, blog, 5, pays + used, P.n - used, len.data() + b0, nullptr);
      for (uint32_t b = b0; b < b0 + k && rc == ZGPU_OK; ++b) used += len[b] == ZGPU_STORED ? 0 : len[b];
    }
    if (rc) { free(pays); return rc; }
    uint64_t t = 0, q = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t raw = (uint32_t)std::min<uint64_t>(B, P.n - (uint64_t)b * B);
      P.meth[b] = len[b] == ZGPU_STORED ? 0 : 1;
      P.clen[b] = P.meth[b] ? len[b] : raw;
      P.coff[b] = t;
      t += P.clen[b];
    }
    P.coff[nb] = t;
    HIPCHK(g, hipHostMalloc((void**)&P.comp, t ? t : 1, hipHostMallocDefault));
    for (uint32_t b = 0; b < nb; ++b) {  // (the payloads come back to back, in block order)
      memcpy(P.comp + P.coff[b], P.meth[b] ? pays + q : P.raw + (uint64_t)b * B, P.clen[b]);
      if (P.meth[b]) q += P.clen[b];
    }
    free(pays);
  }
  // ---- the requests: Poisson arrivals, lognormal sizes, 70 % compress, LZ : random = 9 : 1, levels 1 / 5 / 9
  std::mt19937_64 rng(cfg->seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::normal_distribution<double> N01(0.0, 1.0);
  const double mean_size = cfg->size_median * std::exp(cfg->size_sigma * cfg->size_sigma / 2);
  const double rate = cfg->offered_GBps * 1e9 / mean_size;  // requests per second
  std::vector<SReq> reqs;
  double t = 0;
  uint64_t out_total = 0;
  while (true) {
    t += -std::log(1.0 - U(rng)) / rate;
    if (t > cfg->seconds) break;
    SReq r;
    r.t_arr = t;
    r.op = U(rng) < cfg->compress_share ? 0 : 1;
    const int pool = U(rng) < cfg->random_share ? 1 : 0;
    const double lv = U(rng);
    r.level = lv < 1.0 / 3 ? 1 : lv < 2.0 / 3 ? 5 : 9;
    uint64_t size = (uint64_t)std::llround(std::exp(std::log(cfg->size_median) + cfg->size_sigma * N01(rng)));
    size = std::max<uint64_t>(1, std::min<uint64_t>({size, cfg->size_max, pools[pool].n}));
    Pool& P = pools[pool];
    if (r.op == 0) {
      const uint64_t o = (uint64_t)(U(rng) * (double)(P.n - size + 1));
      r.in = P.raw + std::min(o, P.n - size);
      r.in_len = r.raw_len = size;
    } else {
      const uint32_t nbp = (uint32_t)P.clen.size();
      const uint32_t k = (uint32_t)std::min<uint64_t>((size + B - 1) / B, nbp);
      const uint32_t b0 = (uint32_t)std::min<uint64_t>((uint64_t)(U(rng) * (nbp - k + 1)), nbp - k);
      r.in = P.comp + P.coff[b0];
      r.in_len = P.coff[b0 + k] - P.coff[b0];
      r.raw_len = std::min<uint64_t>((uint64_t)(b0 + k) * B, P.n) - (uint64_t)b0 * B;
      r.nblk = k;
      r.clen = P.clen.data() + b0;
      r.meth = P.meth.data() + b0;
    }
    r.first_blk = (uint32_t)pool;  // (the source pool, until the request is placed in a batch)
    r.ioff = r.op == 0 ? (uint64_t)(r.in - P.raw) : (uint64_t)(r.in - P.comp);
    out_total += r.raw_len;
    reqs.push_back(r);
  }
  std::vector<int> src_pool(reqs.size());
  std::vector<uint64_t> src_off(reqs.size());
  for (size_t i = 0; i < reqs.size(); ++i) { src_pool[i] = (int)reqs[i].first_blk; src_off[i] = reqs[i].ioff; }
  // outputs: pinned, one region per request
  uint8_t* outs = nullptr;
  HIPCHK(g, hipHostMalloc((void**)&outs, out_total ? out_total : 1, hipHostMallocDefault));
  std::vector<uint32_t> out_lens;
  {
    uint64_t nblk = 0;
    for (const SReq& r : reqs) nblk += r.op == 0 ? (r.raw_len + B - 1) / B : 0;
    out_lens.resize(nblk + 1);
    uint64_t o = 0, l = 0;
    for (SReq& r : reqs) {
      r.out = outs + o;
      o += r.raw_len;
      if (r.op == 0) { r.out_len = out_lens.data() + l; l += (r.raw_len + B - 1) / B; }
    }
  }
  // ---- serve
  Server srv;
  if ((rc = srv.init(g, cfg->hbm_budget, cfg->slots, blog))) { srv.release(); return rc; }
  for (Pool& P : pools)
    if (!rc) rc = srv.add_host(P.raw, P.n);
  for (Pool& P : pools)
    if (!rc) rc = srv.add_host(P.comp, P.coff.back() ? P.coff.back() : 1);
  if (!rc) rc = srv.add_host(outs, out_total ? out_total : 1);
  if (rc) { srv.release(); return rc; }
  size_t next = 0;
  const double t0 = now_s();
  for (SReq& r : reqs) r.t_arr += t0;
  while (srv.completed < reqs.size() && rc == ZGPU_OK) {
    const double now = now_s();
    while (next < reqs.size() && reqs[next].t_arr <= now) srv.queue.push_back(&reqs[next++]);
    rc = srv.poll();
  }
  const double t_end = now_s();
  // ---- the report
  std::vector<double> lat;
  lat.reserve(reqs.size());
  uint64_t bytes = 0;
  for (const SReq& r : reqs) {
    lat.push_back((r.t_done - r.t_arr) * 1e3);
    bytes += r.raw_len;
    if (r.op == 0) ++rep->compress_requests; else ++rep->decompress_requests;
  }
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double q) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(q * (double)lat.size()))]; };
  rep->requests = reqs.size();
  rep->bytes = bytes;
  rep->p50_ms = pct(0.50);
  rep->p99_ms = pct(0.99);
  rep->max_ms = lat.empty() ? 0 : lat.back();
  double sum = 0;
  for (double x : lat) sum += x;
  rep->mean_ms = lat.empty() ? 0 : sum / (double)lat.size();
  rep->wall_s = t_end - t0;
  rep->sustained_GBps = rep->wall_s > 0 ? (double)bytes / rep->wall_s / 1e9 : 0;
  rep->batches = srv.batches;
  rep->hbm_bytes = srv.hbm;
  rep->slot_input_bytes = srv.cap_in;
  rep->mean_batch_requests = srv.batches ? (double)srv.batched_reqs / (double)srv.batches : 0;
  // ---- every response round-tripped: decompress outputs against the pool; compress outputs decoded again on the
  // GPU (many requests per call) and compared with their sources
  uint64_t bad = 0;
  struct Pending { size_t req; uint64_t at; };
  std::vector<Pending> pend;
  std::vector<uint32_t> vclen, vrlen, verr;
  std::vector<uint8_t> vmeth, blob, back;
  auto flush = [&]() -> int {
    if (pend.empty()) return ZGPU_OK;
    uint64_t total = 0;
    for (uint32_t x : vrlen) total += x;
    back.resize(total ? total : 1);
    verr.assign(vrlen.size(), 0);
    const int d = zgpu_decompress(g, blob.data(), vclen.data(), vmeth.data(), vrlen.data(), (uint32_t)vrlen.size(),
                                  blog, back.data(), verr.data(), nullptr);
    if (d != ZGPU_OK && d != ZGPU_ERR_DAMAGED) return d;
    for (const Pending& q : pend) {
      const SReq& r = reqs[q.req];
      const uint32_t k0 = (uint32_t)(q.at >> 32);
      uint32_t e = 0;
      for (uint32_t k = 0; k < srv.nblocks(&r); ++k) e |= verr[k0 + k];
      uint64_t o = 0;
      for (uint32_t k = 0; k < k0; ++k) o += vrlen[k];
      if (e || memcmp(back.data() + o, pools[src_pool[q.req]].raw + src_off[q.req], r.raw_len) != 0) ++bad;
    }
    pend.clear(); vclen.clear(); vrlen.clear(); vmeth.clear(); blob.clear();
    return ZGPU_OK;
  };
  for (size_t i = 0; i < reqs.size() && rc == ZGPU_OK; ++i) {
    SReq& r = reqs[i];
    Pool& P = pools[src_pool[i]];
    if (r.op == 1) {  // (src_off: where its stored blocks start in the pool's compressed copy)
      const uint64_t b0 = (uint64_t)(std::lower_bound(P.coff.begin(), P.coff.end(), src_off[i]) - P.coff.begin());
      if (r.err || memcmp(r.out, P.raw + b0 * B, r.raw_len) != 0) ++bad;
      continue;
    }
    if (r.err) { ++bad; continue; }
    const uint8_t* want = P.raw + src_off[i];
    const uint32_t nbk = srv.nblocks(&r);
    pend.push_back({i, (uint64_t)vrlen.size() << 32});
    uint64_t o = 0;
    for (uint32_t k = 0; k < nbk; ++k) {
      const uint32_t rl = (uint32_t)std::min<uint64_t>(B, r.raw_len - (uint64_t)k * B);
      vrlen.push_back(rl);
      if (r.out_len[k] == ZGPU_STORED) {
        vmeth.push_back(0);
        vclen.push_back(rl);
        blob.insert(blob.end(), want + (uint64_t)k * B, want + (uint64_t)k * B + rl);
      } else {
        vmeth.push_back(1);
        vclen.push_back(r.out_len[k]);
        blob.insert(blob.end(), r.out + o, r.out + o + r.out_len[k]);
        o += r.out_len[k];
      }
    }
    if (vrlen.size() >= 4096 || blob.size() >= ((size_t)1 << 30)) rc = flush();
  }
  if (rc == ZGPU_OK) rc = flush();
  rep->mismatches = bad;
  // ---- the sample of responses, for the caller's own checks
  if (cfg->sample_every && rc == ZGPU_OK) {
    uint64_t bo = 0;
    uint32_t lo = 0, ns = 0;
    for (size_t i = 0; i < reqs.size(); i += cfg->sample_every) {
      const SReq& r = reqs[i];
      const Pool& P = pools[src_pool[i]];
      const uint32_t nbk = srv.nblocks(&r);
      const uint64_t len = r.op == 0 ? r.out_bytes : r.raw_len;
      if (ns >= cfg->sample_cap || len > cfg->sample_bytes_cap - bo || (r.op == 0 && nbk > cfg->sample_blk_cap - lo))
        break;
      zgpu_serve_sample_t& sm = cfg->samples[ns++];
      sm.op = (uint32_t)r.op;
      sm.level = (uint32_t)r.level;
      sm.pool = (uint32_t)src_pool[i];
      sm.nblk = nbk;
      sm.src_off = r.op == 0 ? src_off[i]
                             : (uint64_t)(std::lower_bound(P.coff.begin(), P.coff.end(), src_off[i]) - P.coff.begin()) * B;
      sm.raw_len = r.raw_len;
      sm.bytes_off = bo;
      sm.bytes_len = len;
      memcpy(cfg->sample_bytes + bo, r.out, len);
      bo += len;
      sm.blk_len_off = lo;
      if (r.op == 0) {
        memcpy(cfg->sample_blk_len + lo, r.out_len, (size_t)nbk * 4);
        lo += nbk;
      }
      sm.err = r.err;
      sm.latency_ms = (r.t_done - r.t_arr) * 1e3;
    }
    rep->samples = ns;
  }
  // ---- the same requests alone: each one its own batch, compared with the mix
  uint64_t alone_bad = 0;
  const size_t nalone = std::min<size_t>(reqs.size(), (size_t)cfg->alone_checks);
  std::vector<uint8_t> keep;
  std::vector<uint32_t> keep_len;
  for (size_t j = 0; j < nalone && rc == ZGPU_OK; ++j) {
    SReq& r = reqs[(j * 7919) % reqs.size()];
    const uint64_t ob = r.op == 0 ? 0 : r.raw_len;
    keep.assign(r.out, r.out + (r.op == 0 ? r.out_bytes : ob));
    if (r.op == 0) keep_len.assign(r.out_len, r.out_len + srv.nblocks(&r));
    memset(r.out, 0, r.op == 0 ? r.out_bytes : ob);
    srv.queue.push_back(&r);
    while (!srv.idle() && rc == ZGPU_OK) rc = srv.poll();
    if (rc) break;
    const bool same = (r.op == 0 ? r.out_bytes == keep.size() : true) && memcmp(r.out, keep.data(), keep.size()) == 0 &&
                      (r.op == 1 || memcmp(r.out_len, keep_len.data(), keep_len.size() * 4) == 0);
    if (!same) ++alone_bad;
  }
  rep->alone_checked = nalone;
  rep->alone_mismatches = alone_bad;
  srv.release();
  (void)hipHostFree(outs);
  for (Pool& P : pools) { (void)hipHostFree(P.raw); (void)hipHostFree(P.comp); }
  return rc;
}

}  // extern "C"
