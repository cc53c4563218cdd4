// pl_engine.hip — host side of libpilonhip.so: device buffers, the batch
// pipeline, exports, and the C ABI of include/pilon_hip.h.
//
// Layout in HBM (per batch of regions, G positions in total):
//   pile   int32 [PL_NFIELDS][G]   field-major PileUp counters
//   calls  pl_call [G]             BaseCall + disposition, 8 B
//   rec    IndelRec [cap]          insertion/deletion evidence, 32 B each
//   cand   Cand [cap]              change candidates (sparse)
// Resident per contig: the raw reference bytes; per read set: the packed
// reads plus 24 B/read of derived facts (span, alignment end, info word) and
// a 4 B/read prefix count of kept reads.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "pl_kernels.hip"
#include "pl_host.h"

namespace {

thread_local std::string g_err;

pl_status fail(pl_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) return fail(PL_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// Device memory handed out by bumping an offset in a block the library
// holds for the process (the no-allocation entry point's scratch array).
struct Arena {
  uint8_t* base = nullptr;
  size_t size = 0, used = 0;
  bool exhausted = false;
};

struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  Arena* arena = nullptr;   // set: buffers are carved from it, never allocated
  hipError_t ensure(size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (bytes <= cap) return hipSuccess;
    if (arena) {
      const size_t at = (arena->used + 255) & ~(size_t)255;
      if (at > arena->size || bytes > arena->size - at) { arena->exhausted = true; return hipErrorOutOfMemory; }
      p = arena->base + at;
      cap = bytes;
      arena->used = at + bytes;
      return hipSuccess;
    }
    if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; }
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
  void release() { if (p && !arena) (void)hipFree(p); p = nullptr; cap = 0; }
  template <class T> T* as() const { return (T*)p; }
};

struct Contig {
  std::string name;
  std::vector<uint8_t> host;
  uint8_t* dev = nullptr;
  int64_t len = 0;
  // computeGc's buffer value per locus (built on first use), in a device
  // block shared by the contigs built together
  std::shared_ptr<uint8_t> gcv_mem;
  uint8_t* gcv = nullptr;
};

struct ReadSet {
  int32_t contig = -1;
  int64_t n = 0;
  // the BAM kinds the set holds: a frags BAM (GenomeRegion.fragBams non-empty,
  // which turns excludeMotifs off) and a nanopore BAM (GenomeRegion.scala:147,434)
  bool has_frags = true, has_nano = false;
  RSDev d;
  std::vector<void*> mem;
  uint32_t* kp = nullptr;
  int32_t scal[8] = {0};
  int64_t bytes = 0;
  uint32_t nseg = 0;
  // device arenas shared with the other sets of one pl_readset_upload_many
  std::shared_ptr<void> shared[3];
  ~ReadSet() { for (void* p : mem) (void)hipFree(p); }
};

// one region of a batch, by resident read set
struct JobReq {
  int32_t rs, start, stop;
  bool operator==(const JobReq& o) const { return rs == o.rs && start == o.start && stop == o.stop; }
};

// host-owned copy of a pl_read_soa (plain region form)
struct HostReads {
  std::vector<int32_t> pos, tlen;
  std::vector<uint16_t> flag;
  std::vector<uint8_t> mapq;
  std::vector<int64_t> cigar_off{0};
  std::vector<uint32_t> cigar;
  std::vector<int64_t> seq_off{0};
  std::vector<uint8_t> bases2;  // 2-bit codes, one per base here (packed on upload)
  std::vector<int8_t> quals;
  std::vector<int64_t> exo_read, exo_off;
  std::vector<uint8_t> nibc;    // one nibble code per base of exotic reads
};

}  // namespace

// One independent batch pipeline: its stream, events, device buffers and
// host-side tables.  A context has one (the plain and batched forms); a
// server has one per concurrent batch.
struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};
  DBuf rsarr, kparr, jobs, wins, worder, costk, costv, costtk, costtv, pile, calls, st, rec, rkey, rval, rtk, rtv,
      cand, cands, ckey, cval, ctk, ctv, defer, mutated, counters, hist, tiles, strpool, stridx, stroff, gath,
      wide, winwide, heads;
  uint32_t rec_cap = 0, cand_cap = 0, defer_cap = 0, mut_cap = 0, wide_cap = 0;
  std::vector<JobDev> hjobs;
  std::vector<WinDev> hwins;
  std::vector<JobStat> hst;
  std::vector<Cand> hcand;
  std::vector<pl_call> hccall;     // calls of the candidate columns
  std::vector<int64_t> hins_off;   // insertion string of candidate k in hins_pool (-1: none)
  std::string hins_pool;
  int64_t G = 0;
  int32_t W = 0;
  int32_t n_wins = 0;
  bool exported_flags = false;
  double last_ms = 0, last_win_ms = 0;
  uint32_t last_counts[8] = {0};
  int32_t last_persistent = 0;
  int32_t last_block = 0;
  bool last_graph = false;
  EngineArgs args;
  uint32_t* rkey_sorted = nullptr;
  uint32_t* rval_sorted = nullptr;
  // job/window tables on the device describe last_reqs at window W
  std::vector<JobReq> last_reqs;
  bool tables_valid = false;
  uint64_t rs_seen = 0;            // read-set table version on the device
  // captured batch pipeline (pl_batch_opts.use_graph)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gkey;
  EngineArgs gargs;
  // two pinned host buffers for device-to-host hand-offs in chunks (VCF,
  // wig, track values), allocated on first use and kept until the lane is
  // destroyed, so a hand-off allocates no pinned memory
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  // pinned slots for the counters of repeated passes (pl_batch_repeat)
  uint32_t* rep_hc = nullptr;
  // depth of the batch's "jumps" reads before Pass 1, per position, which
  // fragment coverage leaves out (valid for the last batch when set)
  DBuf jdepth;
  bool has_jdepth = false;

  size_t device_bytes() const {
    const DBuf* b[] = {&rsarr, &kparr, &jobs, &wins, &worder, &costk, &costv, &costtk, &costtv, &pile, &calls, &st,
                       &rec, &rkey, &rval, &rtk, &rtv, &cand, &cands, &ckey, &cval, &ctk, &ctv, &defer, &mutated,
                       &counters, &hist, &tiles, &strpool, &stridx, &stroff, &gath, &wide, &winwide, &heads};
    size_t t = 0;
    for (const DBuf* x : b) t += x->cap;
    return t;
  }
  void for_each_buffer(void (*f)(DBuf&, Arena*), Arena* a) {
    DBuf* b[] = {&rsarr, &kparr, &jobs, &wins, &worder, &costk, &costv, &costtk, &costtv, &pile, &calls, &st, &rec,
                 &rkey, &rval, &rtk, &rtv, &cand, &cands, &ckey, &cval, &ctk, &ctv, &defer, &mutated, &counters,
                 &hist, &tiles, &strpool, &stridx, &stroff, &gath, &wide, &winwide, &heads, &jdepth};
    for (DBuf* x : b) f(*x, a);
  }
  // every device buffer of the lane carved from a (no allocation, no free)
  void use_arena(Arena* a) { for_each_buffer([](DBuf& d, Arena* x) { d.release(); d.arena = x; }, a); }
  void free_buffers() {
    for_each_buffer([](DBuf& d, Arena*) { d.release(); }, nullptr);
    tables_valid = false;
    rs_seen = 0;
    rec_cap = cand_cap = defer_cap = mut_cap = wide_cap = 0;
  }
};

// Device memory of pl_batch_run_noalloc: one array per GPU the library runs
// on, held until the process ends.  Host code never uses its name as an
// address: the device address comes from hipGetSymbolAddress, taken once per
// device when a context is created with that device current.
#define PL_NOALLOC_SCRATCH_BYTES ((size_t)192 << 20)
#define PL_NOALLOC_MAX_POSITIONS (1 << 20)
static __device__ __align__(256) uint8_t scratch[PL_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, uint8_t*> g_scratch_ptr;   // device -> address of that device's scratch

// the device address of the current device's scratch array
pl_status scratch_address(int device, uint8_t** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end()) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PL_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    it = g_scratch_ptr.emplace(device, (uint8_t*)ptr).first;
  }
  *out = it->second;
  return PL_OK;
}
}  // namespace

struct pl_ctx {
  pl_config pc;
  KConf cf;
  int device = 0;
  int n_cu = 256;
  std::map<int32_t, Contig> contigs;
  std::vector<std::unique_ptr<ReadSet>> rsets;
  uint64_t rs_version = 1;         // bumped whenever rsets changes
  Lane main;                       // plain and batched forms
  Lane aux;                        // a region's "jumps" reads piled alone
  Lane na;                         // pl_batch_run_noalloc: buffers carved from the device's scratch array
  Arena na_arena;
  std::map<uint64_t, std::pair<int32_t, int32_t>> tuned;   // autotuned (window, block) per batch shape
  int serving = 0;                 // servers open on this context
};

struct pl_region {
  pl_ctx* ctx;
  int32_t contig;
  int32_t start, stop;
  int32_t readset = -1;          // resident set in use
  std::vector<HostReads> batches; // plain add_reads copies
  std::vector<int32_t> types;      // PL_BAM_* of each batch
  bool finished = false;
  pl_region_stats stats;
  std::vector<plh::Change> changes;
  // the read set pl_region_finish built from add_reads: out of the context's
  // table once finished, kept for pl_region_vcf until the region is closed
  std::unique_ptr<ReadSet> temp;
  int32_t temp_id = -1;
  // the "jumps" reads alone (fragment coverage leaves them out), kept the same way
  std::unique_ptr<ReadSet> jtemp;
  int32_t jtemp_id = -1;
  // pl_region_reassemble: bigFixList in the order found, and per attempt
  // logged the region, its outcome (reassemblyFixes) and the log line
  std::vector<plh::Fix> big;
  std::vector<plh::Outcome> outcomes;
  // the loci pl_region_vcf marked deleted for the reassembly's records:
  // Pilon.bed, written after the VCF, counts them as changed
  std::vector<uint8_t> vcf_embed;
};

namespace {
pl_status region_jump_depth(pl_region* r);   // pl_track.hip
}

namespace {

pl_status upload(void** dst, const void* src, size_t bytes, size_t pad, std::vector<void*>& mem, hipStream_t s) {
  HIPCHK(hipMalloc(dst, bytes + pad + 16));
  mem.push_back(*dst);
  if (bytes) HIPCHK(hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s));
  if (pad + 16) HIPCHK(hipMemsetAsync((uint8_t*)*dst + bytes, 0, pad + 16, s));
  return PL_OK;
}

pl_status scan_u32(Lane* L, const uint32_t* in, uint32_t* out, int64_t n) {
  const int64_t tiles = std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
  HIPCHK(L->tiles.ensure(sizeof(uint32_t) * (size_t)tiles));
  k_scan_tiles<<<dim3((unsigned)tiles), dim3(256), 0, L->stream>>>(in, n, L->tiles.as<uint32_t>());
  k_scan_top<<<dim3(1), dim3(1024), 0, L->stream>>>(L->tiles.as<uint32_t>(), tiles);
  k_scan_apply<<<dim3((unsigned)tiles), dim3(256), 0, L->stream>>>(in, n, L->tiles.as<uint32_t>(), out);
  HIPCHK(hipGetLastError());
  return PL_OK;
}

// LSD radix sort of (key, val) pairs whose count is in device memory;
// returns in *rk/*rv the buffers holding the sorted result.
pl_status radix_sort(Lane* L, uint32_t* k, uint32_t* v, uint32_t* tk, uint32_t* tv, const uint32_t* n_dev,
                     uint32_t cap, int bits, uint32_t** rk, uint32_t** rv) {
  const int32_t nb = (int32_t)std::max<int64_t>(1, ((int64_t)cap + RS_TILE - 1) / RS_TILE);
  HIPCHK(L->hist.ensure(sizeof(uint32_t) * ((size_t)256 * nb + 1)));
  uint32_t *ik = k, *iv = v, *ok = tk, *ov = tv;
  const int passes = std::max(1, (bits + 7) / 8);
  for (int p = 0; p < passes; p++) {
    k_rs_hist<<<dim3(nb), dim3(256), 0, L->stream>>>(ik, n_dev, cap, 8 * p, L->hist.as<uint32_t>(), nb);
    pl_status s = scan_u32(L, L->hist.as<uint32_t>(), L->hist.as<uint32_t>(), (int64_t)256 * nb);
    if (s) return s;
    k_rs_scatter<<<dim3(nb), dim3(256), 0, L->stream>>>(ik, iv, ok, ov, n_dev, cap, 8 * p, L->hist.as<uint32_t>(), nb);
    HIPCHK(hipGetLastError());
    std::swap(ik, ok);
    std::swap(iv, ov);
  }
  *rk = ik;
  *rv = iv;
  return PL_OK;
}

int bits_for(uint64_t maxkey) {
  int b = 0;
  while (b < 32 && (maxkey >> b)) b++;
  return std::max(b, 1);
}

void to_kconf(const pl_config& p, KConf& k) {
  k.flank = p.flank; k.min_mq = p.min_mq; k.min_qual = p.min_qual; k.default_qual = p.default_qual;
  k.min_min_depth = p.min_min_depth; k.old_indel = p.old_indel; k.fixamb = p.iupac || p.fix_amb;
  k.nonpf = p.nonpf; k.duplicates = p.duplicates; k.min_depth = p.min_depth;
}

// nibble stream length of a pl_read_soa
int64_t nib_bytes(const pl_read_soa* s) {
  int64_t mx = 0;
  for (int64_t k = 0; k < s->n_exotic; k++) {
    const int64_t r = s->exo_read[k];
    mx = std::max<int64_t>(mx, s->exo_off[k] + (s->seq_off[r + 1] - s->seq_off[r]));
  }
  return (mx + 1) / 2 + 1;
}

// lrt: each read's longReadType (null: none); kinds: RS_FRAGS | RS_NANO
#define RS_FRAGS 1
#define RS_NANO 2
pl_status readset_upload(pl_ctx* c, int32_t contig_id, const pl_read_soa* s, int32_t* out_id,
                         const uint8_t* lrt = nullptr, int kinds = RS_FRAGS) {
  Lane* L = c ? &c->main : nullptr;
  auto it = c->contigs.find(contig_id);
  if (it == c->contigs.end()) return fail(PL_E_INVAL, "contig %d not uploaded", contig_id);
  const int64_t n = s->n_reads;
  if (n < 0) return fail(PL_E_INVAL, "negative read count");
  for (int64_t i = 1; i < n; i++)
    if (s->pos[i] < s->pos[i - 1]) return fail(PL_E_INVAL, "reads are not coordinate-sorted at read %lld", (long long)i);
  for (int64_t k = 1; k < s->n_exotic; k++)
    if (s->exo_read[k] <= s->exo_read[k - 1]) return fail(PL_E_INVAL, "exo_read must be ascending");
  std::unique_ptr<ReadSet> rs(new ReadSet());
  rs->contig = contig_id;
  rs->n = n;
  rs->has_frags = (kinds & RS_FRAGS) != 0;
  rs->has_nano = (kinds & RS_NANO) != 0;
  RSDev& d = rs->d;
  std::memset(&d, 0, sizeof d);
  d.n = n;
  const int64_t ncig = s->cigar_off[n];
  const int64_t nb = s->seq_off[n];
  hipStream_t st = L->stream;
  pl_status e;
  void* p;
#define UP(field, src, bytes, pad)                                  \
  if ((e = upload(&p, (src), (size_t)(bytes), (pad), rs->mem, st))) return e; \
  field = (decltype(field))p;
  UP(d.pos, s->pos, 4 * n, 0);
  UP(d.tlen, s->tlen, 4 * n, 0);
  UP(d.flag, s->flag, 2 * n, 0);
  UP(d.mapq, s->mapq, n, 0);
  UP(d.cigar_off, s->cigar_off, 8 * (n + 1), 0);
  UP(d.cigar, s->cigar, 4 * ncig, 0);
  UP(d.seq_off, s->seq_off, 8 * (n + 1), 0);
  UP(d.bases2, s->bases2, (nb + 3) / 4, 64);
  UP(d.quals, s->quals, nb, 64);
  UP(d.exo_read, s->exo_read, 8 * s->n_exotic, 0);
  UP(d.exo_off, s->exo_off, 8 * s->n_exotic, 0);
  UP(d.nib, s->nib, s->n_exotic ? nib_bytes(s) : 0, 64);
  if (lrt) { UP(d.lrt, lrt, n, 0); }
#undef UP
  d.n_exotic = s->n_exotic;
  d.ref = it->second.dev;
  d.ref_len = it->second.len;
  HIPCHK(hipMalloc(&p, sizeof(RMeta) * (size_t)(n + 1))); rs->mem.push_back(p); d.meta = (RMeta*)p;
  HIPCHK(hipMalloc(&p, 64)); rs->mem.push_back(p); d.scal = (int32_t*)p;
  HIPCHK(hipMalloc(&p, sizeof(FragTail) * (size_t)(n + 1))); rs->mem.push_back(p); d.ftail = (FragTail*)p;
  uint32_t *keepflag, *segcount, *segoff;
  HIPCHK(hipMalloc(&p, 4 * (size_t)(n + 1))); keepflag = (uint32_t*)p;
  HIPCHK(hipMalloc(&p, 4 * (size_t)(n + 1))); segcount = (uint32_t*)p;
  HIPCHK(hipMalloc(&p, 4 * (size_t)(n + 1))); segoff = (uint32_t*)p;
  HIPCHK(hipMalloc(&p, 4 * (size_t)(n + 1))); rs->mem.push_back(p); rs->kp = (uint32_t*)p;
  int64_t ocap = 4096 + n / 256;
  int64_t* outl;
  HIPCHK(hipMalloc(&p, 8 * (size_t)ocap)); rs->mem.push_back(p); outl = (int64_t*)p;
  for (int attempt = 0; attempt < 2; attempt++) {
    HIPCHK(hipMemsetAsync(d.scal, 0, 64, st));
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>(1, (n + 255) / 256), 8192);
    k_prep<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(d, c->cf, keepflag, segcount, d.scal, outl, ocap);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rs->scal, d.scal, 32, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (rs->scal[3] <= ocap) break;
    ocap = rs->scal[3];
    HIPCHK(hipMalloc(&p, 8 * (size_t)ocap)); rs->mem.push_back(p); outl = (int64_t*)p;
  }
  pl_status ss = scan_u32(L, keepflag, rs->kp, n);
  if (ss) return ss;
  // segment table of the multi-op reads
  ss = scan_u32(L, segcount, segoff, n);
  if (ss) return ss;
  uint32_t nseg = 0;
  HIPCHK(hipMemcpyAsync(&nseg, segoff + n, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMalloc(&p, sizeof(Seg) * ((size_t)nseg + 1))); rs->mem.push_back(p);
  d.segs = (const Seg*)p;
  {
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>(1, (n + 255) / 256), 8192);
    k_segfill<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(d, c->cf, segoff, (Seg*)p);
    HIPCHK(hipGetLastError());
  }
  rs->nseg = nseg;
  const int64_t no = rs->scal[3];
  if (no > 0) {
    std::vector<int64_t> h((size_t)no);
    HIPCHK(hipMemcpyAsync(h.data(), outl, 8 * (size_t)no, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::sort(h.begin(), h.end());
    HIPCHK(hipMemcpyAsync(outl, h.data(), 8 * (size_t)no, hipMemcpyHostToDevice, st));
  }
  d.outliers = outl;
  d.n_outlier = no;
  HIPCHK(hipStreamSynchronize(st));
  (void)hipFree(keepflag);
  (void)hipFree(segcount);
  (void)hipFree(segoff);
  rs->bytes = 4 * n + 4 * n + 2 * n + n + 8 * (n + 1) + 4 * ncig + 8 * (n + 1) + (nb + 3) / 4 + nb + 32 * n + 8 * n + 8 * n;
  c->rsets.push_back(std::move(rs));
  c->rs_version++;
  *out_id = (int32_t)c->rsets.size() - 1;
  return PL_OK;
}

// Each set's kept-read prefix and segment offsets from the two scans of
// all sets' flags laid end to end: set i's flags start at B[i], its n + 1
// outputs at OB[i] = B[i] + i; nseg[i] = its segment count.
__global__ void __launch_bounds__(256) k_sets_fix(const uint32_t* ks, const uint32_t* ss, const int64_t* B,
                                                  const int64_t* OB, int32_t n_sets, uint32_t* kp, uint32_t* sego,
                                                  uint32_t* nseg) {
  const int64_t total = OB[n_sets];
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    int32_t lo = 0, hi = n_sets;                 // the last set whose OB <= g
    while (hi - lo > 1) {
      const int32_t mid = (lo + hi) / 2;
      if (OB[mid] <= g) lo = mid; else hi = mid;
    }
    const int64_t b = B[lo], j = g - OB[lo];
    kp[g] = ks[b + j] - ks[b];
    sego[g] = ss[b + j] - ss[b];
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_sets; i += (int64_t)gridDim.x * blockDim.x)
    nseg[i] = ss[B[i + 1]] - ss[B[i]];
}

// The derived tables of read sets whose reads are already on the device
// (each set's d holds its read arrays): per set what readset_upload makes
// after its copies (k_prep's read facts and outlier list, the kept-read
// prefix, the segment table); k_prep and k_segfill per set queued back to
// back on the lane's stream, the scans of all sets at once, a
// synchronisation only where a count from the device is needed, the tables
// in shared arenas.  The sets go into the context's table in order; ids[i]
// = set i's id.
pl_status prep_sets(pl_ctx* c, std::vector<std::unique_ptr<ReadSet>>& rs, const std::vector<int64_t>& ncig,
                    const std::vector<int64_t>& nb, int32_t* ids) {
  Lane* L = &c->main;
  hipStream_t st = L->stream;
  const int32_t n_sets = (int32_t)rs.size();
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  struct Tab { int64_t ocap; size_t meta, ftail, outl; };
  std::vector<Tab> T((size_t)n_sets);
  std::vector<int64_t> Bv((size_t)n_sets + 1, 0), OBv((size_t)n_sets + 1, 0);
  size_t dev = 0;
  for (int32_t i = 0; i < n_sets; i++) {
    const int64_t n = rs[(size_t)i]->n;
    Tab& t = T[(size_t)i];
    t.ocap = 4096 + n / 256;
    t.meta = dev; dev = align(dev + sizeof(RMeta) * (size_t)(n + 1));
    t.ftail = dev; dev = align(dev + sizeof(FragTail) * (size_t)(n + 1));
    t.outl = dev; dev = align(dev + 8 * (size_t)t.ocap);
    Bv[(size_t)i + 1] = Bv[(size_t)i] + n;
    OBv[(size_t)i] = Bv[(size_t)i] + i;
  }
  OBv[(size_t)n_sets] = Bv[(size_t)n_sets] + n_sets;
  const int64_t nall = Bv[(size_t)n_sets], nout = OBv[(size_t)n_sets];
  const size_t kp_at = dev;
  dev = align(dev + 4 * (size_t)nout);
  const size_t scal_at = dev;
  dev = align(dev + 64 * (size_t)std::max(n_sets, 1));
  size_t tmp = 0;
  auto take = [&](size_t bytes) { const size_t a = tmp; tmp = align(tmp + bytes); return a; };
  const size_t keep_at = take(4 * (size_t)(nall + 1)), segc_at = take(4 * (size_t)(nall + 1)),
               ks_at = take(4 * (size_t)(nall + 1)), ss_at = take(4 * (size_t)(nall + 1)),
               sego_at = take(4 * (size_t)nout), b_at = take(8 * Bv.size()), ob_at = take(8 * OBv.size()),
               nseg_at = take(4 * (size_t)std::max(n_sets, 1));
  const size_t v_at = take(sizeof(RSDev) * (size_t)std::max(n_sets, 1)), blk_at = take(8 * ((size_t)n_sets + 1)),
               oc_at = take(8 * (size_t)std::max(n_sets, 1)), ol_at = take(sizeof(int64_t*) * (size_t)std::max(n_sets, 1)),
               sg_at = take(sizeof(Seg*) * (size_t)std::max(n_sets, 1));
  auto dfree = [](void* p) { (void)hipFree(p); };
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, dev));
  std::shared_ptr<void> dev_arena(p, dfree);
  HIPCHK(hipMalloc(&p, std::max<size_t>(tmp, 256)));
  std::unique_ptr<void, decltype(dfree)> tmp_arena(p, dfree);
  uint8_t* D = (uint8_t*)dev_arena.get();
  uint8_t* X = (uint8_t*)tmp_arena.get();
  uint32_t* KP = (uint32_t*)(D + kp_at);
  uint32_t* KEEP = (uint32_t*)(X + keep_at);
  uint32_t* SEGC = (uint32_t*)(X + segc_at);
  uint32_t* SEGO = (uint32_t*)(X + sego_at);
  HIPCHK(hipMemcpyAsync(X + b_at, Bv.data(), 8 * Bv.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(X + ob_at, OBv.data(), 8 * OBv.size(), hipMemcpyHostToDevice, st));
  // the sets' views, then k_prep over all of them in one launch (set i on
  // blocks [blk[i], blk[i + 1]), its flags in its part of the flag arrays)
  std::vector<RSDev> views((size_t)std::max(n_sets, 1));
  std::vector<int64_t> blk((size_t)n_sets + 1, 0), ocaps((size_t)std::max(n_sets, 1));
  std::vector<int64_t*> outls((size_t)std::max(n_sets, 1));
  for (int32_t i = 0; i < n_sets; i++) {
    Tab& t = T[(size_t)i];
    ReadSet& R = *rs[(size_t)i];
    R.shared[1] = dev_arena;
    RSDev& d = R.d;
    const Contig& ct = c->contigs.at(R.contig);
    d.ref = ct.dev;
    d.ref_len = ct.len;
    d.meta = (RMeta*)(D + t.meta);
    d.scal = (int32_t*)(D + scal_at + 64 * (size_t)i);
    d.ftail = (FragTail*)(D + t.ftail);
    R.kp = KP + OBv[(size_t)i];
    views[(size_t)i] = d;
    blk[(size_t)i + 1] = blk[(size_t)i] + std::min<int64_t>(std::max<int64_t>(1, (R.n + 255) / 256), 8192);
    ocaps[(size_t)i] = t.ocap;
    outls[(size_t)i] = (int64_t*)(D + t.outl);
  }
  HIPCHK(hipMemsetAsync(D + scal_at, 0, 64 * (size_t)std::max(n_sets, 1), st));
  if (n_sets > 0) {
    HIPCHK(hipMemcpyAsync(X + v_at, views.data(), sizeof(RSDev) * views.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(X + blk_at, blk.data(), 8 * blk.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(X + oc_at, ocaps.data(), 8 * ocaps.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(X + ol_at, outls.data(), sizeof(int64_t*) * outls.size(), hipMemcpyHostToDevice, st));
    k_prep_sets<<<dim3((unsigned)blk[(size_t)n_sets]), dim3(256), 0, st>>>(
        (const RSDev*)(X + v_at), (const int64_t*)(X + blk_at), n_sets, c->cf, KEEP, SEGC, (const int64_t*)(X + b_at),
        (int64_t* const*)(X + ol_at), (const int64_t*)(X + oc_at));
    HIPCHK(hipGetLastError());
  }
  std::vector<int32_t> scal(16 * (size_t)std::max(n_sets, 1));
  HIPCHK(hipMemcpyAsync(scal.data(), D + scal_at, 64 * (size_t)std::max(n_sets, 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // a set whose outliers overflow their list: a list of the size counted, k_prep again
  std::vector<int64_t*> outl((size_t)n_sets);
  for (int32_t i = 0; i < n_sets; i++) {
    Tab& t = T[(size_t)i];
    ReadSet& R = *rs[(size_t)i];
    outl[(size_t)i] = (int64_t*)(D + t.outl);
    const int32_t cnt = scal[16 * (size_t)i + 3];
    if (cnt > t.ocap) {
      HIPCHK(hipMalloc(&p, 8 * (size_t)cnt));
      R.mem.push_back(p);
      outl[(size_t)i] = (int64_t*)p;
      HIPCHK(hipMemsetAsync(R.d.scal, 0, 64, st));
      const int64_t blocks = std::min<int64_t>(std::max<int64_t>(1, (R.n + 255) / 256), 8192);
      k_prep<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(R.d, c->cf, KEEP + Bv[(size_t)i], SEGC + Bv[(size_t)i],
                                                           R.d.scal, (int64_t*)p, cnt);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&scal[16 * (size_t)i], R.d.scal, 64, hipMemcpyDeviceToHost, st));
    }
  }
  // the kept-read and segment scans of all sets at once, split per set
  HIPCHK(L->tiles.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, (nall + SCAN_TILE - 1) / SCAN_TILE)));
  if (pl_status e = scan_u32(L, KEEP, (uint32_t*)(X + ks_at), nall)) return e;
  if (pl_status e = scan_u32(L, SEGC, (uint32_t*)(X + ss_at), nall)) return e;
  {
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((nout + 255) / 256, 1 << 16));
    k_sets_fix<<<dim3((unsigned)blocks), dim3(256), 0, st>>>((const uint32_t*)(X + ks_at), (const uint32_t*)(X + ss_at),
                                                             (const int64_t*)(X + b_at), (const int64_t*)(X + ob_at),
                                                             n_sets, KP, SEGO, (uint32_t*)(X + nseg_at));
    HIPCHK(hipGetLastError());
  }
  std::vector<uint32_t> nseg((size_t)std::max(n_sets, 1), 0);
  HIPCHK(hipMemcpyAsync(nseg.data(), X + nseg_at, 4 * (size_t)n_sets, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // the segment tables
  size_t segsz = 0;
  std::vector<size_t> segat((size_t)std::max(n_sets, 1));
  for (int32_t i = 0; i < n_sets; i++) { segat[(size_t)i] = segsz; segsz = align(segsz + sizeof(Seg) * ((size_t)nseg[(size_t)i] + 1)); }
  HIPCHK(hipMalloc(&p, std::max<size_t>(segsz, 256)));
  std::shared_ptr<void> seg_arena(p, dfree);
  std::vector<Seg*> segp((size_t)std::max(n_sets, 1));
  for (int32_t i = 0; i < n_sets; i++) {
    ReadSet& R = *rs[(size_t)i];
    R.shared[2] = seg_arena;
    Seg* sg = (Seg*)((uint8_t*)seg_arena.get() + segat[(size_t)i]);
    R.d.segs = sg;
    R.nseg = nseg[(size_t)i];
    segp[(size_t)i] = sg;
  }
  if (n_sets > 0) {                                        // k_segfill over all sets in one launch
    HIPCHK(hipMemcpyAsync(X + sg_at, segp.data(), sizeof(Seg*) * segp.size(), hipMemcpyHostToDevice, st));
    k_segfill_sets<<<dim3((unsigned)blk[(size_t)n_sets]), dim3(256), 0, st>>>(
        (const RSDev*)(X + v_at), (const int64_t*)(X + blk_at), n_sets, c->cf, SEGO, (const int64_t*)(X + ob_at),
        (Seg* const*)(X + sg_at));
    HIPCHK(hipGetLastError());
  }
  // the outlier lists in order: every list to one pinned buffer, sorted
  // there by host threads, and back
  std::vector<int64_t> oat((size_t)n_sets + 1, 0);
  for (int32_t i = 0; i < n_sets; i++) {
    ReadSet& R = *rs[(size_t)i];
    for (int k = 0; k < 8; k++) R.scal[k] = scal[16 * (size_t)i + k];
    oat[(size_t)i + 1] = oat[(size_t)i] + std::max<int64_t>(R.scal[3], 0);
  }
  const int64_t n_out = oat[(size_t)n_sets];
  if (n_out > 0) {
    int64_t* ho = nullptr;
    HIPCHK(hipHostMalloc((void**)&ho, 8 * (size_t)n_out, hipHostMallocDefault));
    std::unique_ptr<int64_t, hipError_t (*)(void*)> hold(ho, hipHostFree);
    for (int32_t i = 0; i < n_sets; i++) {
      const int64_t no = oat[(size_t)i + 1] - oat[(size_t)i];
      if (no > 0)
        HIPCHK(hipMemcpyAsync(ho + oat[(size_t)i], outl[(size_t)i], 8 * (size_t)no, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    std::atomic<int32_t> next(0);
    auto work = [&]() {
      for (int32_t i; (i = next.fetch_add(1)) < n_sets;)
        std::sort(ho + oat[(size_t)i], ho + oat[(size_t)i + 1]);
    };
    std::vector<std::thread> th;
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    try {
      for (int t = 1; t < nt && t < n_sets; t++) th.emplace_back(work);
    } catch (const std::exception&) {
      // fewer threads: the lists left are sorted by the ones running
    }
    work();
    for (auto& t : th) t.join();
    for (int32_t i = 0; i < n_sets; i++) {
      const int64_t no = oat[(size_t)i + 1] - oat[(size_t)i];
      if (no > 0)
        HIPCHK(hipMemcpyAsync(outl[(size_t)i], ho + oat[(size_t)i], 8 * (size_t)no, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
  }
  // the sets into the context's table
  for (int32_t i = 0; i < n_sets; i++) {
    ReadSet& R = *rs[(size_t)i];
    R.d.outliers = outl[(size_t)i];
    R.d.n_outlier = R.scal[3];
    const int64_t n = R.n, nc = ncig[(size_t)i], nbs = nb[(size_t)i];
    R.bytes = 4 * n + 4 * n + 2 * n + n + 8 * (n + 1) + 4 * nc + 8 * (n + 1) + (nbs + 3) / 4 + nbs + 32 * n + 8 * n + 8 * n;
  }
  HIPCHK(hipStreamSynchronize(st));
  for (int32_t i = 0; i < n_sets; i++) {
    c->rsets.push_back(std::move(rs[(size_t)i]));
    ids[i] = (int32_t)c->rsets.size() - 1;
  }
  c->rs_version++;
  return PL_OK;
}

// Many read sets in one pass, each as readset_upload would make it: the
// host arrays of every set packed into one device arena through pinned
// staging buffers (host threads fill one while the other is copied), the
// per-set passes (k_prep, the two scans, k_segfill) queued back to back on
// the lane's stream with one synchronisation between the passes that need
// a count from the device, and the derived tables in two more arenas.
pl_status readsets_upload_many(pl_ctx* c, int32_t n_sets, const int32_t* contig_ids, const pl_read_soa* const* sets,
                               const int32_t* bam_types, int32_t* out_ids) {
  Lane* L = &c->main;
  hipStream_t st = L->stream;
  struct Field { const void* src; size_t bytes, pad; size_t at; };
  struct Plan {
    int64_t n, ncig, nb;
    std::vector<Field> f;              // pos tlen flag mapq cigar_off cigar seq_off bases2 quals exo_read exo_off nib [lrt]
    std::vector<uint8_t> lrt;
  };
  std::vector<Plan> P((size_t)n_sets);
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t up = 0;
  for (int32_t i = 0; i < n_sets; i++) {
    const pl_read_soa* s = sets[i];
    if (!s) return fail(PL_E_INVAL, "read set %d is null", i);
    if (c->contigs.find(contig_ids[i]) == c->contigs.end()) return fail(PL_E_INVAL, "contig %d not uploaded", contig_ids[i]);
    const int32_t bt = bam_types ? bam_types[i] : PL_BAM_FRAGS;
    if (bt != PL_BAM_FRAGS && bt != PL_BAM_UNPAIRED && bt != PL_BAM_NANOPORE && bt != PL_BAM_PACBIO)
      return fail(PL_E_INVAL, "read set %d: bam type %d of a resident read set is not PL_BAM_FRAGS, UNPAIRED, NANOPORE or "
                  "PACBIO", i, bt);
    const bool lr = bt == PL_BAM_NANOPORE || bt == PL_BAM_PACBIO;
    if (lr && !c->pc.longread) return fail(PL_E_INVAL, "a long-read BAM needs config.longread (Pilon.longread)");
    const int64_t n = s->n_reads;
    if (n < 0) return fail(PL_E_INVAL, "read set %d: negative read count", i);
    for (int64_t k = 1; k < n; k++)
      if (s->pos[k] < s->pos[k - 1])
        return fail(PL_E_INVAL, "read set %d: reads are not coordinate-sorted at read %lld", i, (long long)k);
    for (int64_t k = 1; k < s->n_exotic; k++)
      if (s->exo_read[k] <= s->exo_read[k - 1]) return fail(PL_E_INVAL, "read set %d: exo_read must be ascending", i);
    Plan& q = P[(size_t)i];
    q.n = n;
    q.ncig = s->cigar_off[n];
    q.nb = s->seq_off[n];
    if (lr) q.lrt.assign((size_t)std::max<int64_t>(n, 1), bt == PL_BAM_NANOPORE ? PL_LR_NANOPORE : PL_LR_PACBIO);
    q.f = {{s->pos, (size_t)(4 * n), 0, 0}, {s->tlen, (size_t)(4 * n), 0, 0}, {s->flag, (size_t)(2 * n), 0, 0},
           {s->mapq, (size_t)n, 0, 0}, {s->cigar_off, (size_t)(8 * (n + 1)), 0, 0}, {s->cigar, (size_t)(4 * q.ncig), 0, 0},
           {s->seq_off, (size_t)(8 * (n + 1)), 0, 0}, {s->bases2, (size_t)((q.nb + 3) / 4), 64, 0},
           {s->quals, (size_t)q.nb, 64, 0}, {s->exo_read, (size_t)(8 * s->n_exotic), 0, 0},
           {s->exo_off, (size_t)(8 * s->n_exotic), 0, 0}, {s->nib, (size_t)(s->n_exotic ? nib_bytes(s) : 0), 64, 0}};
    if (lr) q.f.push_back({q.lrt.data(), (size_t)n, 0, 0});
    for (Field& f : q.f) { f.at = up; up = align(up + f.bytes + f.pad + 16); }
  }
  HIPCHK(hipSetDevice(c->device));
  auto dfree = [](void* p) { (void)hipFree(p); };
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(up, 256)));
  std::shared_ptr<void> up_arena(p, dfree);
  uint8_t* U = (uint8_t*)up_arena.get();
  // the host arrays through two pinned staging buffers, packed by host
  // threads into the arena's layout (padding zeroed)
  {
    const size_t chunk = std::min<size_t>(align(up), (size_t)128 << 20);
    void* sb[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    pl_status e = PL_OK;
    for (int b = 0; b < 2 && !e; b++) {
      if (hipHostMalloc(&sb[b], chunk, hipHostMallocDefault) != hipSuccess) e = fail(PL_E_OOM, "pinned staging of %zu bytes", chunk);
      else if (hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess) e = fail(PL_E_DEVICE, "event");
    }
    std::vector<const Field*> all;                          // in arena order
    for (const Plan& q : P) for (const Field& f : q.f) all.push_back(&f);
    auto field_end = [&](size_t k) { return k + 1 < all.size() ? all[k + 1]->at : up; };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    int b = 0;
    for (size_t a0 = 0; a0 < up && !e; a0 += chunk, b ^= 1) {
      const size_t a1 = std::min(up, a0 + chunk);
      if (hipEventSynchronize(ev[b]) != hipSuccess) { e = fail(PL_E_DEVICE, "staging event"); break; }
      uint8_t* S = (uint8_t*)sb[b];
      // nt pieces of [a0, a1), each over the fields (with their padding and
      // alignment gap, zeroed) that meet it, taken by whichever thread is free
      const int pieces = nt;
      std::atomic<int> next(0);
      auto work = [&]() {
        for (int w; (w = next.fetch_add(1)) < pieces;) {
          const size_t lo = a0 + (a1 - a0) * (size_t)w / (size_t)pieces;
          const size_t hi = a0 + (a1 - a0) * (size_t)(w + 1) / (size_t)pieces;
          size_t k = (size_t)(std::upper_bound(all.begin(), all.end(), lo,
                                               [](size_t x, const Field* f) { return x < f->at; }) - all.begin());
          k = k > 0 ? k - 1 : 0;
          for (; k < all.size() && all[k]->at < hi; k++) {
            const Field& f = *all[k];
            const size_t x0 = std::max(lo, f.at), x1 = std::min(hi, field_end(k));
            if (x0 >= x1) continue;
            const size_t d1 = std::min(x1, f.at + f.bytes);
            if (d1 > x0) std::memcpy(S + (x0 - a0), (const uint8_t*)f.src + (x0 - f.at), d1 - x0);
            const size_t z0 = std::max(x0, f.at + f.bytes);
            if (x1 > z0) std::memset(S + (z0 - a0), 0, x1 - z0);
          }
        }
      };
      std::vector<std::thread> th;
      try {
        for (int t = 1; t < nt; t++) th.emplace_back(work);
      } catch (const std::exception&) {
        // fewer threads: the pieces left are taken by the ones running
      }
      work();
      for (auto& t : th) t.join();
      if (hipMemcpyAsync(U + a0, S, a1 - a0, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipEventRecord(ev[b], st) != hipSuccess)
        e = fail(PL_E_DEVICE, "staged copy");
    }
    (void)hipStreamSynchronize(st);
    for (int k = 0; k < 2; k++) {
      if (ev[k]) (void)hipEventDestroy(ev[k]);
      if (sb[k]) (void)hipHostFree(sb[k]);
    }
    if (e) return e;
  }
  // the sets' device views
  std::vector<std::unique_ptr<ReadSet>> rs((size_t)n_sets);
  std::vector<int64_t> ncig((size_t)n_sets), nb((size_t)n_sets);
  for (int32_t i = 0; i < n_sets; i++) {
    Plan& q = P[(size_t)i];
    const int32_t bt = bam_types ? bam_types[i] : PL_BAM_FRAGS;
    rs[(size_t)i].reset(new ReadSet());
    ReadSet& R = *rs[(size_t)i];
    R.contig = contig_ids[i];
    R.n = q.n;
    R.has_frags = bt == PL_BAM_FRAGS;
    R.has_nano = bt == PL_BAM_NANOPORE;
    R.shared[0] = up_arena;
    RSDev& d = R.d;
    std::memset(&d, 0, sizeof d);
    d.n = q.n;
    d.pos = (const int32_t*)(U + q.f[0].at); d.tlen = (const int32_t*)(U + q.f[1].at);
    d.flag = (const uint16_t*)(U + q.f[2].at); d.mapq = (const uint8_t*)(U + q.f[3].at);
    d.cigar_off = (const int64_t*)(U + q.f[4].at); d.cigar = (const uint32_t*)(U + q.f[5].at);
    d.seq_off = (const int64_t*)(U + q.f[6].at); d.bases2 = (const uint8_t*)(U + q.f[7].at);
    d.quals = (const int8_t*)(U + q.f[8].at); d.exo_read = (const int64_t*)(U + q.f[9].at);
    d.exo_off = (const int64_t*)(U + q.f[10].at); d.nib = (const uint8_t*)(U + q.f[11].at);
    if (q.f.size() > 12) d.lrt = (const uint8_t*)(U + q.f[12].at);
    d.n_exotic = sets[i]->n_exotic;
    ncig[(size_t)i] = q.ncig;
    nb[(size_t)i] = q.nb;
  }
  return prep_sets(c, rs, ncig, nb, out_ids);
}

// Everything a captured graph bakes in: the kernel arguments (buffer
// pointers, capacities, sizes) and the launch shapes.  A replay is valid
// only while all of it is unchanged.
struct GraphKey {
  EngineArgs A;
  int32_t W, block, persistent, grid;
  uint64_t lds;
  const void* bufs[8];
};

void drop_graph(Lane* L) {
  if (L->gexec) (void)hipGraphExecDestroy(L->gexec);
  if (L->graph) (void)hipGraphDestroy(L->graph);
  L->gexec = nullptr;
  L->graph = nullptr;
}

pl_status lane_init(Lane* L) {
  HIPCHK(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
  for (int i = 0; i < 4; i++) HIPCHK(hipEventCreate(&L->ev[i]));
  return PL_OK;
}

void lane_destroy(Lane* L) {
  if (L->stream) (void)hipStreamSynchronize(L->stream);
  drop_graph(L);
  L->free_buffers();
  for (int k = 0; k < 2; k++) {
    if (L->stage_ev[k]) (void)hipEventDestroy(L->stage_ev[k]);
    if (L->stage[k]) (void)hipHostFree(L->stage[k]);
    L->stage_ev[k] = nullptr;
    L->stage[k] = nullptr;
  }
  if (L->rep_hc) (void)hipHostFree(L->rep_hc);
  L->rep_hc = nullptr;
  for (int i = 0; i < 4; i++) if (L->ev[i]) (void)hipEventDestroy(L->ev[i]);
  if (L->stream) (void)hipStreamDestroy(L->stream);
  L->stream = nullptr;
  for (int i = 0; i < 4; i++) L->ev[i] = nullptr;
}

// the device pipeline of one batch, enqueued on the lane's stream; with
// timed set, events bracket the window kernel (not while capturing)
pl_status enqueue_batch(Lane* L, EngineArgs& A, int persistent, int block, int grid, size_t lds, bool timed) {
  hipStream_t st = L->stream;
  uint32_t* cnt = L->counters.as<uint32_t>();
  const int32_t n_jobs = A.n_jobs, n_wins = A.n_wins;
  const int64_t G = A.G;
  const int64_t words = (int64_t)n_jobs * (sizeof(JobStat) / 4);
  k_init_stats<<<dim3((unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (words + 255) / 256))), dim3(256), 0, st>>>(A.st, n_jobs, cnt, 16);
  k_init_stats2<<<dim3((n_jobs + 255) / 256), dim3(256), 0, st>>>(A.st, n_jobs);
  k_plan<<<dim3((n_wins + 255) / 256), dim3(256), 0, st>>>(A.rs, A.jobs, L->wins.as<WinDev>(), n_wins, L->costk.as<uint32_t>());
  HIPCHK(hipGetLastError());
  A.worder = L->worder.as<int32_t>();
  if (persistent) {
    k_iota<<<dim3((unsigned)std::min<int64_t>(1024, (n_wins + 255) / 256)), dim3(256), 0, st>>>(L->costv.as<uint32_t>(), n_wins, cnt + 8);
    uint32_t *sk, *sv;
    pl_status s = radix_sort(L, L->costk.as<uint32_t>(), L->costv.as<uint32_t>(), L->costtk.as<uint32_t>(),
                             L->costtv.as<uint32_t>(), cnt + 8, (uint32_t)n_wins, 16, &sk, &sv);
    if (s) return s;
    A.worder = (const int32_t*)sv;
  }
  k_win_heads<<<dim3((n_wins + 255) / 256), dim3(256), 0, st>>>(A.jobs, A.wins, persistent ? A.worder : nullptr, n_wins,
                                                                 (WinHead*)A.heads);
  k_jobprep<<<dim3(n_jobs), dim3(64), 0, st>>>(A.rs, (const uint32_t* const*)L->kparr.p, A.jobs, A.st);
  HIPCHK(hipGetLastError());
  if (timed) HIPCHK(hipEventRecord(L->ev[2], st));
  if (block == 256) k_window<256><<<dim3(grid), dim3(256), lds, st>>>(A, persistent);
  else if (block == 512) k_window<512><<<dim3(grid), dim3(512), lds, st>>>(A, persistent);
  else k_window<1024><<<dim3(grid), dim3(1024), lds, st>>>(A, persistent);
  HIPCHK(hipGetLastError());
  if (timed) HIPCHK(hipEventRecord(L->ev[3], st));
  // evidence records: sort by (position, kind), then the deferred columns
  const uint32_t gb = (uint32_t)std::min<int64_t>(4096, std::max<int64_t>(1, ((int64_t)A.rec_cap + 255) / 256));
  k_rec_keys<<<dim3(gb), dim3(256), 0, st>>>(A.rec, A.rec_n, A.rec_cap, L->rkey.as<uint32_t>(), L->rval.as<uint32_t>());
  uint32_t *rk, *rv;
  pl_status s = radix_sort(L, L->rkey.as<uint32_t>(), L->rval.as<uint32_t>(), L->rtk.as<uint32_t>(), L->rtv.as<uint32_t>(),
                           A.rec_n, A.rec_cap, bits_for(2 * (uint64_t)G), &rk, &rv);
  if (s) return s;
  A.rkey = rk; A.rval = rv;
  L->rkey_sorted = rk; L->rval_sorted = rv;
  const uint32_t gd = (uint32_t)std::min<int64_t>(2048, std::max<int64_t>(1, ((int64_t)A.defer_cap + 3) / 4));
  k_indel_pass<<<dim3(gd), dim3(256), 0, st>>>(A, A.defer, A.defer_n, A.defer_cap, 0);
  HIPCHK(hipGetLastError());
  // candidates by position, then Pass 1 per region
  const uint32_t gc = (uint32_t)std::min<int64_t>(1024, std::max<int64_t>(1, ((int64_t)A.cand_cap + 255) / 256));
  k_cand_keys<<<dim3(gc), dim3(256), 0, st>>>(A.cand, A.cand_n, A.cand_cap, L->ckey.as<uint32_t>(), L->cval.as<uint32_t>());
  uint32_t *ck, *cv;
  s = radix_sort(L, L->ckey.as<uint32_t>(), L->cval.as<uint32_t>(), L->ctk.as<uint32_t>(), L->ctv.as<uint32_t>(),
                 A.cand_n, A.cand_cap, bits_for((uint64_t)G), &ck, &cv);
  if (s) return s;
  k_cand_gather<<<dim3(gc), dim3(256), 0, st>>>(A.cand, A.cand_n, A.cand_cap, cv, L->cands.as<Cand>());
  k_finalize<<<dim3(n_jobs), dim3(64), 0, st>>>(A, L->cands.as<Cand>(), A.cand_n, A.cand_cap, L->cands.as<Cand>());
  k_confirm_exact<<<dim3((unsigned)std::min<int32_t>(n_wins, 1024)), dim3(256), 0, st>>>(A);
  const uint32_t gm = (uint32_t)std::min<int64_t>(2048, std::max<int64_t>(1, ((int64_t)A.mutated_cap + 3) / 4));
  k_indel_pass<<<dim3(gm), dim3(256), 0, st>>>(A, A.mutated, A.mutated_n, A.mutated_cap, 1);
  HIPCHK(hipGetLastError());
  return PL_OK;
}

// LDS the window kernel needs for a window of W positions (dynamic part)
size_t window_lds(int32_t W) { return PL_WIN_LDS_PER_POS * (size_t)W + 1024; }
// ... and per workgroup in total, with the per-wave run tables
size_t window_lds_total(int32_t W, int block) {
  return window_lds(W) + 32 * (size_t)block + 1024 + sizeof(IndelRec) * PL_WIN_RECS;
}

pl_status run_jobs(pl_ctx* c, Lane* L, const std::vector<JobReq>& reqs, int32_t W, int persistent, int block, bool use_graph) {
  if (reqs.empty()) return fail(PL_E_INVAL, "empty batch");
  if (W < 64 || W % 64 || W > 2048) return fail(PL_E_INVAL, "window %d must be a multiple of 64 in [64, 2048]", W);
  if (block != 256 && block != 512 && block != 1024) return fail(PL_E_INVAL, "block %d must be 256, 512 or 1024", block);
  if (window_lds_total(W, block) > 160 * 1024)
    return fail(PL_E_INVAL, "window %d with block %d needs more than the 160 KB of LDS of a CU", W, block);
  L->exported_flags = false;
  L->has_jdepth = false;
  hipStream_t st = L->stream;
  // ---- jobs and windows (kept on the device while the batch repeats)
  const bool same = L->tables_valid && L->W == W && reqs == L->last_reqs;
  if (!same) {
    L->tables_valid = false;
    L->W = W;
    L->hjobs.resize(reqs.size());
    L->hwins.clear();
    int64_t G = 0;
    for (size_t j = 0; j < reqs.size(); j++) {
      const JobReq& q = reqs[j];
      if (q.rs < 0 || q.rs >= (int32_t)c->rsets.size() || !c->rsets[q.rs]) return fail(PL_E_INVAL, "bad read set %d", q.rs);
      const Contig& ct = c->contigs.at(c->rsets[q.rs]->contig);
      if (q.start < 1 || q.stop < q.start || q.stop > ct.len)
        return fail(PL_E_INVAL, "region %d-%d outside contig of length %lld", q.start, q.stop, (long long)ct.len);
      JobDev& J = L->hjobs[j];
      std::memset(&J, 0, sizeof J);
      J.rs = q.rs; J.start = q.start; J.stop = q.stop;
      J.qs = (int32_t)std::max<int64_t>((int64_t)q.start - 10000, 0);
      J.qe = (int32_t)std::min<int64_t>((int64_t)q.stop + 10000, ct.len);
      J.size = q.stop - q.start + 1;
      J.out = G;
      J.win0 = (int32_t)L->hwins.size();
      for (int32_t p0 = 0; p0 < J.size; p0 += W) {
        WinDev w;
        std::memset(&w, 0, sizeof w);
        w.job = (int32_t)j; w.p0 = p0; w.p1 = std::min<int32_t>(J.size, p0 + W);
        L->hwins.push_back(w);
      }
      J.nwin = (int32_t)L->hwins.size() - J.win0;
      G += J.size;
    }
    if (G >= (int64_t)1 << 31) return fail(PL_E_INVAL, "batch of %lld positions exceeds 2^31", (long long)G);
    L->G = G;
    L->n_wins = (int32_t)L->hwins.size();
  }
  const int64_t G = L->G;
  const int32_t n_jobs = (int32_t)reqs.size();
  const int32_t n_wins = L->n_wins;
  // ---- resident read-set table
  if (L->rs_seen != c->rs_version) {
    std::vector<RSDev> rsd(c->rsets.size());
    std::vector<const uint32_t*> kps(c->rsets.size(), nullptr);
    for (size_t i = 0; i < c->rsets.size(); i++) if (c->rsets[i]) { rsd[i] = c->rsets[i]->d; kps[i] = c->rsets[i]->kp; }
    HIPCHK(L->rsarr.ensure(sizeof(RSDev) * rsd.size()));
    HIPCHK(L->kparr.ensure(sizeof(void*) * kps.size()));
    HIPCHK(hipMemcpyAsync(L->rsarr.p, rsd.data(), sizeof(RSDev) * rsd.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(L->kparr.p, kps.data(), sizeof(void*) * kps.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    L->rs_seen = c->rs_version;
  }
  // ---- capacities (grown and re-run on overflow)
  if (!L->rec_cap) { L->rec_cap = 1u << 18; L->cand_cap = 1u << 16; L->defer_cap = 1u << 16; L->mut_cap = 1u << 16; }
  if (!L->wide_cap) L->wide_cap = 1u << 12;
  L->rec_cap = std::max<uint32_t>(L->rec_cap, (uint32_t)std::min<int64_t>(G / 16 + 65536, 1ll << 30));
  L->cand_cap = std::max<uint32_t>(L->cand_cap, (uint32_t)std::min<int64_t>(G / 256 + 16384, 1ll << 30));
  L->defer_cap = std::max<uint32_t>(L->defer_cap, (uint32_t)std::min<int64_t>(G / 256 + 16384, 1ll << 30));
  L->mut_cap = std::max<uint32_t>(L->mut_cap, (uint32_t)std::min<int64_t>(G / 256 + 16384, 1ll << 30));
  HIPCHK(L->jobs.ensure(sizeof(JobDev) * n_jobs));
  HIPCHK(L->wins.ensure(sizeof(WinDev) * n_wins));
  HIPCHK(L->worder.ensure(4 * (size_t)n_wins));
  HIPCHK(L->winwide.ensure(8 * (size_t)n_wins));
  HIPCHK(L->heads.ensure(sizeof(WinHead) * (size_t)n_wins));
  HIPCHK(L->costk.ensure(4 * (size_t)n_wins));
  HIPCHK(L->costv.ensure(4 * (size_t)n_wins));
  HIPCHK(L->costtk.ensure(4 * (size_t)n_wins));
  HIPCHK(L->costtv.ensure(4 * (size_t)n_wins));
  HIPCHK(L->pile.ensure(4 * (size_t)PL_NFIELDS * G));
  HIPCHK(L->calls.ensure(sizeof(pl_call) * (size_t)G));
  HIPCHK(L->st.ensure(sizeof(JobStat) * n_jobs));
  HIPCHK(L->counters.ensure(64));
  if (!same) {
    HIPCHK(hipMemcpyAsync(L->jobs.p, L->hjobs.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(L->wins.p, L->hwins.data(), sizeof(WinDev) * n_wins, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    L->last_reqs = reqs;
    L->tables_valid = true;
  }
  const size_t lds = window_lds(W);
  const void* kfn = block == 256 ? (const void*)k_window<256> : block == 512 ? (const void*)k_window<512>
                                                                             : (const void*)k_window<1024>;
  HIPCHK(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int grid = n_wins;
  if (persistent) {
    // workgroups per CU: LDS (dynamic + static run tables) and 32 waves
    const int by_lds = (int)(160 * 1024 / window_lds_total(W, block));
    const int by_waves = 32 / (block / 64);
    const int per_cu = std::max(1, std::min(by_lds, by_waves));
    grid = std::min(n_wins, c->n_cu * per_cu);
  }
  for (int attempt = 0; attempt < 6; attempt++) {
    HIPCHK(L->rec.ensure(sizeof(IndelRec) * (size_t)L->rec_cap));
    HIPCHK(L->rkey.ensure(4 * (size_t)L->rec_cap)); HIPCHK(L->rval.ensure(4 * (size_t)L->rec_cap));
    HIPCHK(L->rtk.ensure(4 * (size_t)L->rec_cap)); HIPCHK(L->rtv.ensure(4 * (size_t)L->rec_cap));
    HIPCHK(L->cand.ensure(sizeof(Cand) * (size_t)L->cand_cap));
    HIPCHK(L->cands.ensure(sizeof(Cand) * (size_t)L->cand_cap));
    HIPCHK(L->ckey.ensure(4 * (size_t)L->cand_cap)); HIPCHK(L->cval.ensure(4 * (size_t)L->cand_cap));
    HIPCHK(L->ctk.ensure(4 * (size_t)L->cand_cap)); HIPCHK(L->ctv.ensure(4 * (size_t)L->cand_cap));
    HIPCHK(L->defer.ensure(4 * (size_t)L->defer_cap));
    HIPCHK(L->mutated.ensure(4 * (size_t)L->mut_cap));
    HIPCHK(L->wide.ensure(sizeof(WideCol) * (size_t)L->wide_cap));
    // sort scratch for the largest sort, allocated before any capture
    {
      const uint32_t mx = std::max<uint32_t>(std::max<uint32_t>(L->rec_cap, L->cand_cap), (uint32_t)n_wins);
      const int64_t nb = std::max<int64_t>(1, ((int64_t)mx + RS_TILE - 1) / RS_TILE);
      HIPCHK(L->hist.ensure(sizeof(uint32_t) * ((size_t)256 * nb + 1)));
      HIPCHK(L->tiles.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, (256 * nb + SCAN_TILE - 1) / SCAN_TILE)));
    }
    uint32_t* cnt = L->counters.as<uint32_t>();
    EngineArgs& A = L->args;
    std::memset(&A, 0, sizeof A);
    A.cf = c->cf;
    A.rs = L->rsarr.as<RSDev>();
    A.jobs = L->jobs.as<JobDev>();
    A.n_jobs = n_jobs;
    A.wins = L->wins.as<WinDev>();
    A.n_wins = n_wins;
    A.W = W;
    A.queue = (int32_t*)(cnt + 5);
    A.deep = cnt + 7;
    A.worder = L->worder.as<int32_t>();
    A.heads = L->heads.as<WinHead>();
    A.pile = L->pile.as<int32_t>();
    A.calls = L->calls.as<pl_call>();
    A.G = G;
    A.st = L->st.as<JobStat>();
    A.rec = L->rec.as<IndelRec>(); A.rec_n = cnt + 0; A.rec_cap = L->rec_cap;
    A.cand = L->cand.as<Cand>(); A.cand_n = cnt + 1; A.cand_cap = L->cand_cap;
    A.defer = L->defer.as<uint32_t>(); A.defer_n = cnt + 2; A.defer_cap = L->defer_cap;
    A.mutated = L->mutated.as<uint32_t>(); A.mutated_n = cnt + 3; A.mutated_cap = L->mut_cap;
    A.overflow = (int32_t*)(cnt + 4);
    A.wide = L->wide.as<WideCol>(); A.wide_n = cnt + 6; A.wide_cap = L->wide_cap;
    A.win_wide = L->winwide.as<int32_t>();

    if (use_graph) {
      GraphKey key;
      std::memset(&key, 0, sizeof key);
      key.A = A;
      key.W = W; key.block = block; key.persistent = persistent; key.grid = grid; key.lds = lds;
      key.bufs[0] = L->hist.p; key.bufs[1] = L->tiles.p; key.bufs[2] = L->rsarr.p; key.bufs[3] = L->kparr.p;
      key.bufs[4] = L->cands.p; key.bufs[5] = L->costtk.p; key.bufs[6] = L->rtk.p; key.bufs[7] = L->ctk.p;
      if (!L->gexec || L->gkey.size() != sizeof key || std::memcmp(L->gkey.data(), &key, sizeof key) != 0) {
        drop_graph(L);
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        EngineArgs Acap = A;
        pl_status s = enqueue_batch(L, Acap, persistent, block, grid, lds, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (s) { if (g) (void)hipGraphDestroy(g); return s; }
        HIPCHK(e);
        L->graph = g;
        HIPCHK(hipGraphInstantiate(&L->gexec, g, nullptr, nullptr, 0));
        L->gkey.assign((const uint8_t*)&key, (const uint8_t*)&key + sizeof key);
        L->gargs = Acap;
      }
      A = L->gargs;
      HIPCHK(hipEventRecord(L->ev[0], st));
      HIPCHK(hipGraphLaunch(L->gexec, st));
      HIPCHK(hipEventRecord(L->ev[1], st));
    } else {
      HIPCHK(hipEventRecord(L->ev[0], st));
      pl_status s = enqueue_batch(L, A, persistent, block, grid, lds, true);
      if (s) return s;
      HIPCHK(hipEventRecord(L->ev[1], st));
    }
    L->rkey_sorted = A.rkey; L->rval_sorted = A.rval;
    uint32_t hc[8];
    HIPCHK(hipMemcpyAsync(hc, cnt, 32, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0, wms = -1.0f;
    HIPCHK(hipEventElapsedTime(&ms, L->ev[0], L->ev[1]));
    if (!use_graph) HIPCHK(hipEventElapsedTime(&wms, L->ev[2], L->ev[3]));
    L->last_ms = ms;
    L->last_win_ms = wms;
    L->last_graph = use_graph;
    std::memcpy(L->last_counts, hc, sizeof hc);
    L->last_persistent = persistent;
    L->last_block = block;
    bool grow = false;
    if (hc[0] > L->rec_cap) { L->rec_cap = (uint32_t)std::min<uint64_t>(2ull * hc[0], 1u << 30); grow = true; }
    if (hc[1] > L->cand_cap) { L->cand_cap = (uint32_t)std::min<uint64_t>(2ull * hc[1], 1u << 30); grow = true; }
    if (hc[2] > L->defer_cap) { L->defer_cap = (uint32_t)std::min<uint64_t>(2ull * hc[2], 1u << 30); grow = true; }
    if (hc[3] > L->mut_cap) { L->mut_cap = (uint32_t)std::min<uint64_t>(2ull * hc[3], 1u << 30); grow = true; }
    if (hc[6] > L->wide_cap) { L->wide_cap = (uint32_t)std::min<uint64_t>(2ull * hc[6], 1u << 30); grow = true; }
    if (!grow) {
      if (hc[4]) return fail(PL_E_DEVICE, "capacity overflow flag set without an overfull counter");
      return PL_OK;
    }
  }
  return fail(PL_E_OOM, "evidence buffers kept overflowing");
}

// Batch shape the tuned launch configuration is cached under: log2 buckets
// of the positions, the region count and the read density.
uint64_t shape_key(pl_ctx* c, const std::vector<JobReq>& q, int persistent) {
  double G = 0, reads = 0;
  for (const JobReq& j : q) {
    const ReadSet* rs = (j.rs >= 0 && j.rs < (int32_t)c->rsets.size()) ? c->rsets[j.rs].get() : nullptr;
    const double size = (double)j.stop - j.start + 1;
    G += size;
    if (rs) {
      const double len = (double)std::max<int64_t>(1, c->contigs.at(rs->contig).len);
      reads += (double)rs->n * size / len;
    }
  }
  auto lg = [](double x) { return (uint64_t)std::floor(std::log2(std::max(1.0, x))); };
  return lg(G) | (lg((double)q.size()) << 8) | (lg(16.0 * reads / std::max(1.0, G)) << 16) | ((uint64_t)persistent << 24);
}

// Picks (window, block) for a batch by timing every candidate that fits the
// LDS of a CU on the batch itself (each trial is a complete, valid run) and
// caches the fastest per batch shape.  A caller-fixed window or block is
// kept and only the other one is searched.
pl_status autotune(pl_ctx* c, Lane* L, const std::vector<JobReq>& q, int persistent, int32_t* W, int32_t* block) {
  const uint64_t key = shape_key(c, q, persistent);
  auto it = c->tuned.find(key);
  if (it == c->tuned.end()) {
    static const int32_t Ws[] = {256, 384, 512, 640, 768, 1024};
    static const int32_t Bs[] = {256, 512};
    double best = 1e30;
    std::pair<int32_t, int32_t> pick(PL_DEFAULT_WINDOW, PL_DEFAULT_BLOCK);
    for (int32_t w : Ws) {
      if (*W && w != *W) continue;
      for (int32_t b : Bs) {
        if (*block && b != *block) continue;
        if (window_lds_total(w, b) > 160 * 1024) continue;
        pl_status s = run_jobs(c, L, q, w, persistent, b, false);
        if (s) return s;
        if (L->last_ms < best) { best = L->last_ms; pick = {w, b}; }
      }
    }
    it = c->tuned.emplace(key, pick).first;
  }
  if (!*W) *W = it->second.first;
  if (!*block) *block = it->second.second;
  return PL_OK;
}

// copy stats (and candidates) of the last batch to the host
// Stats, candidates and what the change lists need of them (the calls of
// the candidate columns, the insertion strings) for every job of the last
// batch: a fixed number of launches and copies whatever the batch size.
pl_status fetch_batch_meta(Lane* L) {
  const size_t nj = L->hjobs.size();
  L->hst.resize(nj);
  HIPCHK(hipMemcpyAsync(L->hst.data(), L->st.p, sizeof(JobStat) * nj, hipMemcpyDeviceToHost, L->stream));
  uint32_t cn = 0;
  HIPCHK(hipMemcpyAsync(&cn, L->counters.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  cn = std::min(cn, L->cand_cap);
  L->hcand.resize(cn);
  L->hccall.resize(cn);
  L->hins_off.assign(cn, -1);
  L->hins_pool.clear();
  if (cn) {
    HIPCHK(hipMemcpyAsync(L->hcand.data(), L->cands.p, sizeof(Cand) * cn, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(L->gath.ensure(sizeof(pl_call) * cn));
    k_cand_calls<<<dim3((unsigned)std::min<uint32_t>(1024, (cn + 255) / 256)), dim3(256), 0, L->stream>>>(
        L->cands.as<Cand>(), cn, L->calls.as<pl_call>(), L->gath.as<pl_call>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(L->hccall.data(), L->gath.p, sizeof(pl_call) * cn, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    // insertion strings of every INS candidate, decoded on the device at once
    std::vector<int64_t> idx, off(1, 0);
    for (uint32_t k = 0; k < cn; k++)
      if (L->hcand[k].kind == PL_D_INS) {
        L->hins_off[k] = off.back();
        idx.push_back(k);
        off.push_back(off.back() + L->hcand[k].len);
      }
    if (!idx.empty()) {
      HIPCHK(L->stridx.ensure(8 * idx.size()));
      HIPCHK(L->stroff.ensure(8 * off.size()));
      HIPCHK(L->strpool.ensure((size_t)off.back() + 16));
      HIPCHK(hipMemcpyAsync(L->stridx.p, idx.data(), 8 * idx.size(), hipMemcpyHostToDevice, L->stream));
      HIPCHK(hipMemcpyAsync(L->stroff.p, off.data(), 8 * off.size(), hipMemcpyHostToDevice, L->stream));
      k_ins_strings<<<dim3((unsigned)((idx.size() + 63) / 64)), dim3(64), 0, L->stream>>>(
          L->args, L->cands.as<Cand>(), L->stridx.as<int64_t>(), L->stroff.as<int64_t>(), (int32_t)idx.size(),
          L->strpool.as<uint8_t>());
      HIPCHK(hipGetLastError());
      L->hins_pool.resize((size_t)off.back());
      HIPCHK(hipMemcpyAsync(&L->hins_pool[0], L->strpool.p, (size_t)off.back(), hipMemcpyDeviceToHost, L->stream));
      HIPCHK(hipStreamSynchronize(L->stream));
    }
  }
  return PL_OK;
}

pl_status job_status(Lane* L, size_t j) {
  const JobStat& S = L->hst[j];
  if (S.err_read_inval != INT64_MAX && S.err_read_inval <= S.err_read_edge)
    return fail(PL_E_INVAL, "malformed read %lld (CIGAR and SEQ disagree or bad op)", (long long)S.err_read_inval);
  if (S.err_read_edge != INT64_MAX)
    return fail(PL_E_EDGE, "read %lld: Pilon 1.24 throws here (indel normalised before region start, or clip without SEQ)",
                (long long)S.err_read_edge);
  if (S.overflow) return fail(PL_E_OVERFLOW, "a column holds %d or more bases and deletions", PL_DEPTH_LIMIT);
  return PL_OK;
}

// stats + change list of job j of the last batch (needs fetch_batch_meta)
pl_status job_changes(pl_ctx* c, Lane* L, size_t j, pl_region_stats* out, std::vector<plh::Change>* changes) {
  const JobDev& J = L->hjobs[j];
  const JobStat& S = L->hst[j];
  const Contig& ct = c->contigs.at(c->rsets[J.rs]->contig);
  std::vector<plh::Change> ch;
  for (int32_t k = S.cand_lo; k < S.cand_hi && k < (int32_t)L->hcand.size(); k++) {
    const Cand& cd = L->hcand[k];
    if (!cd.kind) continue;
    plh::Change x;
    x.loc = J.start + (int32_t)(cd.gpos - J.out);
    x.kind = cd.kind;
    x.ref = to_upper(ct.host[x.loc - 1]);
    x.base = L->hccall[k].base;
    x.alt = L->hccall[k].alt;
    if (cd.kind == PL_D_DEL) x.indel.assign((const char*)ct.host.data() + x.loc - 1, (size_t)cd.len);
    if (cd.kind == PL_D_INS) x.indel = L->hins_pool.substr((size_t)L->hins_off[k], (size_t)cd.len);
    ch.push_back(x);
  }
  const ReadSet& rs = *c->rsets[J.rs];
  if (c->pc.longread && !rs.has_frags) plh::exclude_motifs(ch, ct.host.data(), J.start, J.stop, rs.has_nano);
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->read_count = S.read_count;
    out->base_count = S.base_count;
    out->coverage = round_div_l(S.base_count, J.size);
    out->min_depth = S.min_depth;
    out->n_confirmed = S.n_confirmed;
    out->n_unknown_ops = (int32_t)S.unknown_ops;
    out->aligned_bases = S.aligned;
    for (const plh::Change& x : ch) {
      if (x.kind == PL_D_SNP) out->snps++;
      else if (x.kind == PL_D_AMB) { if (c->pc.fix_snps && !c->pc.longread) out->amb++; }
      else if (x.kind == PL_D_INS) { out->ins++; out->ins_bases += (int32_t)x.indel.size(); }
      else if (x.kind == PL_D_DEL) { out->dels++; out->del_bases += (int32_t)x.indel.size(); }
    }
  }
  if (changes) *changes = std::move(ch);
  return PL_OK;
}

// pileup / final calls of job j of the last batch into host buffers
pl_status job_export(Lane* L, size_t j, int32_t* pile, pl_call* calls) {
  const JobDev& J = L->hjobs[j];
  if (calls && !L->exported_flags) {
    k_final_flags<<<dim3(L->n_wins), dim3(256), 0, L->stream>>>(L->args);
    HIPCHK(hipGetLastError());
    L->exported_flags = true;
  }
  if (pile)
    for (int f = 0; f < PL_NFIELDS; f++)
      HIPCHK(hipMemcpyAsync(pile + (size_t)f * J.size, L->pile.as<int32_t>() + (size_t)f * L->G + J.out,
                            4 * (size_t)J.size, hipMemcpyDeviceToHost, L->stream));
  if (calls)
    HIPCHK(hipMemcpyAsync(calls, L->calls.as<pl_call>() + J.out, sizeof(pl_call) * (size_t)J.size,
                          hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return PL_OK;
}

void append_reads(HostReads& h, const pl_read_soa* s) {
  const int64_t n = s->n_reads;
  const int64_t r0 = (int64_t)h.pos.size();
  h.pos.insert(h.pos.end(), s->pos, s->pos + n);
  h.tlen.insert(h.tlen.end(), s->tlen, s->tlen + n);
  h.flag.insert(h.flag.end(), s->flag, s->flag + n);
  h.mapq.insert(h.mapq.end(), s->mapq, s->mapq + n);
  const int64_t c0 = h.cigar_off.back(), s0 = h.seq_off.back();
  for (int64_t i = 0; i < n; i++) {
    h.cigar_off.push_back(c0 + s->cigar_off[i + 1]);
    h.seq_off.push_back(s0 + s->seq_off[i + 1]);
  }
  h.cigar.insert(h.cigar.end(), s->cigar, s->cigar + s->cigar_off[n]);
  const int64_t nb = s->seq_off[n];
  h.quals.insert(h.quals.end(), s->quals, s->quals + nb);
  for (int64_t q = 0; q < nb; q++) h.bases2.push_back((s->bases2[q >> 2] >> (2 * (q & 3))) & 3);
  for (int64_t k = 0; k < s->n_exotic; k++) {
    const int64_t r = s->exo_read[k];
    const int64_t L = s->seq_off[r + 1] - s->seq_off[r];
    h.exo_read.push_back(r0 + r);
    h.exo_off.push_back((int64_t)h.nibc.size());
    for (int64_t q = 0; q < L; q++) {
      const int64_t x = s->exo_off[k] + q;
      const uint8_t b = s->nib[x >> 1];
      h.nibc.push_back((x & 1) ? (b & 15) : (b >> 4));
    }
  }
}

// merge the plain-form batches (of PL_BAM_* types) into one coordinate-sorted set and upload it
pl_status upload_host_batches(pl_ctx* c, int32_t contig, std::vector<HostReads>& bs, const std::vector<int32_t>& types,
                              int32_t* id) {
  struct Ref { int32_t b; int64_t r; };
  std::vector<Ref> order;
  for (int32_t b = 0; b < (int32_t)bs.size(); b++)
    for (int64_t r = 0; r < (int64_t)bs[b].pos.size(); r++) order.push_back({b, r});
  std::stable_sort(order.begin(), order.end(), [&](const Ref& x, const Ref& y) { return bs[x.b].pos[x.r] < bs[y.b].pos[y.r]; });
  const int64_t n = (int64_t)order.size();
  std::vector<int32_t> pos(n), tlen(n);
  std::vector<uint16_t> flag(n);
  std::vector<uint8_t> mapq(n);
  std::vector<int64_t> cigar_off(n + 1, 0), seq_off(n + 1, 0), exo_read, exo_off;
  std::vector<uint32_t> cigar;
  std::vector<int8_t> quals;
  std::vector<uint8_t> codes, nibc, lrt(n);
  bool any_lr = false;
  int kinds = 0;
  for (int32_t t : types) kinds |= (t == PL_BAM_FRAGS ? RS_FRAGS : 0) | (t == PL_BAM_NANOPORE ? RS_NANO : 0);
  for (int64_t i = 0; i < n; i++) {
    const int32_t t = types[order[i].b];
    lrt[i] = t == PL_BAM_NANOPORE ? PL_LR_NANOPORE : t == PL_BAM_PACBIO ? PL_LR_PACBIO : 0;
    any_lr = any_lr || lrt[i];
    const HostReads& h = bs[order[i].b];
    const int64_t r = order[i].r;
    pos[i] = h.pos[r]; tlen[i] = h.tlen[r]; flag[i] = h.flag[r]; mapq[i] = h.mapq[r];
    cigar.insert(cigar.end(), h.cigar.begin() + h.cigar_off[r], h.cigar.begin() + h.cigar_off[r + 1]);
    quals.insert(quals.end(), h.quals.begin() + h.seq_off[r], h.quals.begin() + h.seq_off[r + 1]);
    codes.insert(codes.end(), h.bases2.begin() + h.seq_off[r], h.bases2.begin() + h.seq_off[r + 1]);
    cigar_off[i + 1] = (int64_t)cigar.size();
    seq_off[i + 1] = (int64_t)quals.size();
    auto it = std::lower_bound(h.exo_read.begin(), h.exo_read.end(), r);
    if (it != h.exo_read.end() && *it == r) {
      const int64_t k = it - h.exo_read.begin();
      exo_read.push_back(i);
      exo_off.push_back((int64_t)nibc.size());
      nibc.insert(nibc.end(), h.nibc.begin() + h.exo_off[k], h.nibc.begin() + h.exo_off[k] + (h.seq_off[r + 1] - h.seq_off[r]));
    }
  }
  std::vector<uint8_t> bases2((codes.size() + 3) / 4 + 8, 0), nib(nibc.size() / 2 + 8, 0);
  for (size_t q = 0; q < codes.size(); q++) bases2[q >> 2] |= (uint8_t)(codes[q] << (2 * (q & 3)));
  for (size_t q = 0; q < nibc.size(); q++) nib[q >> 1] |= (uint8_t)((q & 1) ? nibc[q] : (nibc[q] << 4));
  pl_read_soa s;
  std::memset(&s, 0, sizeof s);
  s.n_reads = n;
  s.pos = pos.data(); s.tlen = tlen.data(); s.flag = flag.data(); s.mapq = mapq.data();
  s.cigar_off = cigar_off.data(); s.cigar = cigar.data(); s.seq_off = seq_off.data();
  s.bases2 = bases2.data(); s.quals = quals.data();
  s.n_exotic = (int64_t)exo_read.size(); s.exo_read = exo_read.data(); s.exo_off = exo_off.data(); s.nib = nib.data();
  return readset_upload(c, contig, &s, id, any_lr ? lrt.data() : nullptr, kinds);
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

void pl_config_default(pl_config* c) {
  std::memset(c, 0, sizeof *c);
  c->flank = 10; c->min_mq = 0; c->min_qual = 0; c->default_qual = 10; c->min_depth = 0.1;
  c->min_min_depth = 5; c->fix_snps = 1; c->fix_indels = 1;
}

const char* pl_last_error(void) { return g_err.c_str(); }

pl_status pl_ctx_create(const pl_config* cfg, int device, pl_ctx** out) {
  if (!cfg || !out) return fail(PL_E_INVAL, "null argument");
  if (cfg->flank < 0) return fail(PL_E_INVAL, "flank must be >= 0");
  if (cfg->default_qual < -128 || cfg->default_qual > 127) return fail(PL_E_INVAL, "defaultQual is a Byte");
  if (cfg->min_min_depth < 0) return fail(PL_E_INVAL, "minMinDepth must be >= 0");
  if (!std::isfinite(cfg->min_depth)) return fail(PL_E_INVAL, "minDepth must be a finite number");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(PL_E_DEVICE, "no HIP device %d (have %d)", device, n);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  std::unique_ptr<pl_ctx> c(new pl_ctx());
  c->pc = *cfg;
  to_kconf(*cfg, c->cf);
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  pl_status st = lane_init(&c->main);
  if (!st) st = lane_init(&c->aux);
  if (!st) st = lane_init(&c->na);
  if (!st) st = scratch_address(device, &c->na_arena.base);
  if (st) { lane_destroy(&c->main); lane_destroy(&c->aux); lane_destroy(&c->na); return st; }
  c->na_arena.size = PL_NOALLOC_SCRATCH_BYTES;
  c->na.use_arena(&c->na_arena);
  *out = c.release();
  return PL_OK;
}

void pl_ctx_destroy(pl_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  lane_destroy(&c->main);
  lane_destroy(&c->aux);
  lane_destroy(&c->na);
  c->rsets.clear();
  for (auto& kv : c->contigs) if (kv.second.dev) (void)hipFree(kv.second.dev);
  c->contigs.clear();
  delete c;
}

pl_status pl_ref_upload(pl_ctx* c, int32_t contig_id, const char* name, const uint8_t* ascii, int64_t len) {
  Lane* L = c ? &c->main : nullptr;
  if (!c || !ascii || len <= 0) return fail(PL_E_INVAL, "bad reference");
  if (len >= ((int64_t)1 << 31) - 1) return fail(PL_E_INVAL, "contig longer than 2^31");
  if (c->serving) return fail(PL_E_BUSY, "a server is open on this context: its contigs are fixed");
  HIPCHK(hipSetDevice(c->device));
  Contig& ct = c->contigs[contig_id];
  if (ct.dev) { HIPCHK(hipStreamSynchronize(L->stream)); HIPCHK(hipFree(ct.dev)); ct.dev = nullptr; }
  ct.gcv_mem.reset();
  ct.gcv = nullptr;
  ct.name = name ? name : "";
  ct.host.assign(ascii, ascii + len);
  ct.len = len;
  HIPCHK(hipMalloc((void**)&ct.dev, (size_t)len + 64));
  HIPCHK(hipMemcpyAsync(ct.dev, ascii, (size_t)len, hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemsetAsync(ct.dev + len, 0, 64, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return PL_OK;
}

pl_status pl_readset_upload_typed(pl_ctx* c, int32_t contig_id, const pl_read_soa* reads, int32_t bam_type,
                                  int32_t* id) {
  if (!c || !reads || !id) return fail(PL_E_INVAL, "null argument");
  if (bam_type != PL_BAM_FRAGS && bam_type != PL_BAM_UNPAIRED && bam_type != PL_BAM_NANOPORE &&
      bam_type != PL_BAM_PACBIO)
    return fail(PL_E_INVAL, "bam type %d of a resident read set is not PL_BAM_FRAGS, UNPAIRED, NANOPORE or PACBIO "
                "(a jumps BAM goes through pl_region_add_reads_typed)", bam_type);
  const bool lr = bam_type == PL_BAM_NANOPORE || bam_type == PL_BAM_PACBIO;
  if (lr && !c->pc.longread) return fail(PL_E_INVAL, "a long-read BAM needs config.longread (Pilon.longread)");
  if (c->serving) return fail(PL_E_BUSY, "a server is open on this context: its read sets are fixed");
  HIPCHK(hipSetDevice(c->device));
  std::vector<uint8_t> lrt;
  if (lr) lrt.assign((size_t)std::max<int64_t>(reads->n_reads, 1), bam_type == PL_BAM_NANOPORE ? PL_LR_NANOPORE : PL_LR_PACBIO);
  return readset_upload(c, contig_id, reads, id, lr ? lrt.data() : nullptr,
                        (bam_type == PL_BAM_FRAGS ? RS_FRAGS : 0) | (bam_type == PL_BAM_NANOPORE ? RS_NANO : 0));
}

pl_status pl_readset_upload(pl_ctx* c, int32_t contig_id, const pl_read_soa* reads, int32_t* id) {
  return pl_readset_upload_typed(c, contig_id, reads, PL_BAM_FRAGS, id);
}

pl_status pl_ctx_synchronize(pl_ctx* c) {
  if (!c) return fail(PL_E_INVAL, "null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  return PL_OK;
}

pl_status pl_readset_upload_many(pl_ctx* c, int32_t n_sets, const int32_t* contig_ids, const pl_read_soa* const* reads,
                                 const int32_t* bam_types, int32_t* ids) {
  if (!c || n_sets < 0 || (n_sets > 0 && (!contig_ids || !reads || !ids))) return fail(PL_E_INVAL, "null argument");
  if (c->serving) return fail(PL_E_BUSY, "a server is open on this context: its read sets are fixed");
  if (n_sets == 0) return PL_OK;
  try {
    return readsets_upload_many(c, n_sets, contig_ids, reads, bam_types, ids);
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed");
  }
}

pl_status pl_region_open(pl_ctx* c, int32_t contig_id, int32_t start, int32_t stop, pl_region** out) {
  if (!c || !out) return fail(PL_E_INVAL, "null argument");
  auto it = c->contigs.find(contig_id);
  if (it == c->contigs.end()) return fail(PL_E_INVAL, "contig %d not uploaded", contig_id);
  if (start < 1 || stop < start || stop > it->second.len)
    return fail(PL_E_INVAL, "GenomeRegion stop point must be within contig (%d-%d of %lld)", start, stop, (long long)it->second.len);
  pl_region* r = new pl_region();
  r->ctx = c; r->contig = contig_id; r->start = start; r->stop = stop;
  std::memset(&r->stats, 0, sizeof r->stats);
  *out = r;
  return PL_OK;
}

pl_status pl_region_add_reads_typed(pl_region* r, const pl_read_soa* reads, int32_t bam_type) {
  if (!r || !reads) return fail(PL_E_INVAL, "null argument");
  if (bam_type < PL_BAM_FRAGS || bam_type > PL_BAM_PACBIO)
    return fail(PL_E_INVAL, "bam type %d is not a PL_BAM_* type", bam_type);
  if ((bam_type == PL_BAM_NANOPORE || bam_type == PL_BAM_PACBIO) && !r->ctx->pc.longread)
    return fail(PL_E_INVAL, "a long-read BAM needs config.longread (Pilon.longread)");
  if (r->finished) return fail(PL_E_INVAL, "region already finished");
  if (r->readset >= 0) return fail(PL_E_INVAL, "region already uses a resident read set");
  for (int64_t i = 1; i < reads->n_reads; i++)
    if (reads->pos[i] < reads->pos[i - 1]) return fail(PL_E_INVAL, "reads must be coordinate-sorted within a call");
  r->batches.emplace_back();   // one batch per BAM; merged by position at finish
  r->types.push_back(bam_type);
  if (reads->n_reads > 0) append_reads(r->batches.back(), reads);
  return PL_OK;
}

pl_status pl_region_add_reads(pl_region* r, const pl_read_soa* reads) {
  return pl_region_add_reads_typed(r, reads, PL_BAM_FRAGS);
}

pl_status pl_region_use_readset(pl_region* r, int32_t id) {
  if (!r) return fail(PL_E_INVAL, "null argument");
  if (!r->batches.empty() || r->finished) return fail(PL_E_INVAL, "region already has reads");
  if (id < 0 || id >= (int32_t)r->ctx->rsets.size() || !r->ctx->rsets[id] || r->ctx->rsets[id]->contig != r->contig)
    return fail(PL_E_INVAL, "read set %d is not a set of this contig", id);
  r->readset = id;
  return PL_OK;
}

pl_status pl_region_finish(pl_region* r, int32_t* pile, pl_call* calls, pl_region_stats* stats) {
  if (!r) return fail(PL_E_INVAL, "null argument");
  pl_ctx* c = r->ctx;
  Lane* L = &c->main;
  HIPCHK(hipSetDevice(c->device));
  int32_t id = r->readset;
  bool temp = false;
  bool jumps = false;
  if (id < 0) {
    if (c->serving) return fail(PL_E_BUSY, "a server is open on this context: use a resident read set");
    if (r->batches.empty()) { r->batches.emplace_back(); r->types.push_back(PL_BAM_FRAGS); }
    for (int32_t t : r->types) jumps = jumps || t == PL_BAM_JUMPS;
    if (jumps) {
      // the jumps BAMs alone, piled beside the region for fragment coverage
      std::vector<HostReads> jb;
      std::vector<int32_t> jt;
      for (size_t k = 0; k < r->batches.size(); k++)
        if (r->types[k] == PL_BAM_JUMPS) { jb.push_back(r->batches[k]); jt.push_back(PL_BAM_JUMPS); }
      int32_t jid = -1;
      pl_status s = upload_host_batches(c, r->contig, jb, jt, &jid);
      if (s) return s;
      r->jtemp = std::move(c->rsets[jid]);
      r->jtemp_id = jid;
      c->rs_version++;
    }
    pl_status s = upload_host_batches(c, r->contig, r->batches, r->types, &id);
    if (s) return s;
    temp = true;
    r->batches.clear();
  }
  std::vector<JobReq> q{{id, r->start, r->stop}};
  pl_status s = run_jobs(c, L, q, PL_DEFAULT_WINDOW, 0, PL_DEFAULT_BLOCK, false);
  if (!s) s = fetch_batch_meta(L);
  if (!s) s = job_status(L, 0);
  if (!s) s = job_changes(c, L, 0, &r->stats, &r->changes);
  if (!s) s = job_export(L, 0, pile, calls);
  if (temp) {
    r->temp = std::move(c->rsets[id]);
    r->temp_id = id;
    c->rs_version++;
  }
  if (!s && jumps) s = region_jump_depth(r);
  if (s) return s;
  if (stats) *stats = r->stats;
  r->finished = true;
  return PL_OK;
}

pl_status pl_region_fix(pl_region* r, const char* new_name, int64_t offset, char** changes_text, int64_t* changes_len,
                        uint8_t** bases, int64_t* bases_len) {
  if (!r || !r->finished) return fail(PL_E_INVAL, "region not finished");
  const Contig& ct = r->ctx->contigs[r->contig];
  std::string rb((const char*)ct.host.data() + r->start - 1, (size_t)(r->stop - r->start + 1));
  std::string nn = new_name ? new_name : plh::pilon_name(ct.name);
  // bigFixList holds the reassembly's fixes last found first
  const std::vector<plh::Fix> big(r->big.rbegin(), r->big.rend());
  plh::FixResult fr = plh::fix_region(r->ctx->pc, ct.name, r->start, rb, r->changes, nn, offset, big);
  if (changes_text) {
    *changes_text = (char*)malloc(fr.changes.size() + 1);
    std::memcpy(*changes_text, fr.changes.c_str(), fr.changes.size() + 1);
  }
  if (changes_len) *changes_len = (int64_t)fr.changes.size();
  if (bases) {
    *bases = (uint8_t*)malloc(fr.bases.size() + 1);
    std::memcpy(*bases, fr.bases.data(), fr.bases.size());
  }
  if (bases_len) *bases_len = (int64_t)fr.bases.size();
  return PL_OK;
}

void pl_region_close(pl_region* r) { delete r; }
void pl_free(void* p) { free(p); }

}  // extern "C"

namespace {

pl_status batch_requests(pl_ctx* c, const pl_region_desc* regions, int32_t n, std::vector<JobReq>* q) {
  q->resize((size_t)n);
  for (int32_t i = 0; i < n; i++) {
    const pl_region_desc& d = regions[i];
    if (d.readset_id < 0 || d.readset_id >= (int32_t)c->rsets.size() || !c->rsets[d.readset_id] ||
        c->rsets[d.readset_id]->contig != d.contig_id)
      return fail(PL_E_INVAL, "region %d: read set %d is not a set of contig %d", i, d.readset_id, d.contig_id);
    (*q)[i] = {d.readset_id, d.start, d.stop};
  }
  return PL_OK;
}

// stats and host copies of the lane's last batch (pl_batch_run's outputs)
pl_status batch_outputs(pl_ctx* c, Lane* L, int32_t n, int32_t* pile, pl_call* calls, pl_region_stats* stats) {
  pl_status s = fetch_batch_meta(L);
  if (s) return s;
  for (int32_t i = 0; i < n; i++) if ((s = job_status(L, i))) return s;
  if (stats)
    for (int32_t i = 0; i < n; i++) if ((s = job_changes(c, L, i, &stats[i], nullptr))) return s;
  if (pile || calls) {
    int64_t off = 0;
    for (int32_t i = 0; i < n; i++) {
      const int64_t sz = L->hjobs[i].size;
      std::vector<int32_t> tmp;
      if (pile) tmp.resize((size_t)PL_NFIELDS * sz);
      if ((s = job_export(L, i, pile ? tmp.data() : nullptr, calls ? calls + off : nullptr))) return s;
      if (pile)
        for (int f = 0; f < PL_NFIELDS; f++)
          std::memcpy(pile + (size_t)f * L->G + off, tmp.data() + (size_t)f * sz, 4 * (size_t)sz);
      off += sz;
    }
  }
  return PL_OK;
}

}  // namespace

extern "C" {

pl_status pl_batch_run(pl_ctx* c, const pl_region_desc* regions, int32_t n, const pl_batch_opts* opts, int32_t* pile,
                       pl_call* calls, pl_region_stats* stats) {
  if (!c || !regions || n <= 0) return fail(PL_E_INVAL, "empty batch");
  Lane* L = &c->main;
  HIPCHK(hipSetDevice(c->device));
  std::vector<JobReq> q;
  pl_status s = batch_requests(c, regions, n, &q);
  if (s) return s;
  int32_t W = opts ? opts->window : 0;
  int32_t block = opts ? opts->block : 0;
  const int persistent = opts ? opts->persistent : 0;
  const bool use_graph = opts && opts->use_graph;
  if (W == 0 || block == 0) {
    if ((s = autotune(c, L, q, persistent, &W, &block))) return s;
  }
  if ((s = run_jobs(c, L, q, W, persistent, block, use_graph))) return s;
  return batch_outputs(c, L, n, pile, calls, stats);
}

pl_status pl_batch_run_noalloc(pl_ctx* c, const pl_region_desc* regions, int32_t n, const pl_batch_opts* opts,
                               int32_t* pile, pl_call* calls, pl_region_stats* stats) {
  if (!c || !regions || n <= 0) return fail(PL_E_INVAL, "empty batch");
  Lane* L = &c->na;
  HIPCHK(hipSetDevice(c->device));
  std::vector<JobReq> q;
  pl_status s = batch_requests(c, regions, n, &q);
  if (s) return s;
  int64_t G = 0;
  for (int32_t i = 0; i < n; i++) G += std::max<int64_t>(0, (int64_t)regions[i].stop - regions[i].start + 1);
  if (G > PL_NOALLOC_MAX_POSITIONS)
    return fail(PL_E_INVAL, "batch of %lld positions exceeds the no-allocation limit of %d", (long long)G,
                PL_NOALLOC_MAX_POSITIONS);
  if (opts && opts->use_graph) return fail(PL_E_INVAL, "the no-allocation form does not capture graphs");
  const int32_t W = opts && opts->window ? opts->window : PL_DEFAULT_WINDOW;
  const int32_t block = opts && opts->block ? opts->block : PL_DEFAULT_BLOCK;
  const int persistent = opts ? opts->persistent : 0;
  // every buffer of this call is carved afresh from the start of the array;
  // what must start at zero is cleared by the pipeline's own kernels
  L->free_buffers();
  c->na_arena.used = 0;
  c->na_arena.exhausted = false;
  s = run_jobs(c, L, q, W, persistent, block, false);
  if (!s) s = batch_outputs(c, L, n, pile, calls, stats);
  if (s && c->na_arena.exhausted)
    return fail(PL_E_OOM, "the batch's device buffers exceed the %zu-byte scratch array", PL_NOALLOC_SCRATCH_BYTES);
  return s;
}

pl_status pl_batch_repeat(pl_ctx* c, int32_t n_passes) {
  if (!c || n_passes < 0) return fail(PL_E_INVAL, "null argument or negative pass count");
  if (c->serving) return fail(PL_E_BUSY, "a server is open on this context");
  Lane* L = &c->main;
  if (!L->tables_valid || !L->last_graph || !L->gexec)
    return fail(PL_E_INVAL, "the last batch was not run with use_graph (nothing to repeat)");
  HIPCHK(hipSetDevice(c->device));
  constexpr int kSlots = 64;   // passes in flight between two checks
  if (!L->rep_hc) {
    void* p = nullptr;
    if (hipHostMalloc(&p, sizeof(uint32_t) * 8 * kSlots, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PL_E_OOM, "pinned host allocation failed");
    }
    L->rep_hc = (uint32_t*)p;
  }
  hipStream_t st = L->stream;
  const uint32_t* cnt = L->counters.as<uint32_t>();
  for (int32_t done = 0; done < n_passes;) {
    const int32_t m = std::min<int32_t>(kSlots, n_passes - done);
    HIPCHK(hipEventRecord(L->ev[0], st));
    for (int32_t i = 0; i < m; i++) {
      HIPCHK(hipGraphLaunch(L->gexec, st));
      HIPCHK(hipMemcpyAsync(L->rep_hc + 8 * i, cnt, 32, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipEventRecord(L->ev[1], st));
    HIPCHK(hipStreamSynchronize(st));
    for (int32_t i = 0; i < m; i++) {
      const uint32_t* hc = L->rep_hc + 8 * i;
      if (hc[0] > L->rec_cap || hc[1] > L->cand_cap || hc[2] > L->defer_cap || hc[3] > L->mut_cap ||
          hc[6] > L->wide_cap || hc[4])
        return fail(PL_E_OOM, "pass %d of the repeated batch overflowed its buffers: run it with pl_batch_run",
                    done + i);
    }
    std::memcpy(L->last_counts, L->rep_hc + 8 * (m - 1), sizeof L->last_counts);
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, L->ev[0], L->ev[1]));
    L->last_ms = ms / (float)m;
    done += m;
  }
  return PL_OK;
}

pl_status pl_last_batch_info(pl_ctx* c, pl_batch_info* o) {
  Lane* L = c ? &c->main : nullptr;
  if (!c || !o) return fail(PL_E_INVAL, "null argument");
  std::memset(o, 0, sizeof *o);
  o->positions = L->G;
  o->n_regions = (int32_t)L->hjobs.size();
  o->n_windows = L->n_wins;
  o->n_records = L->last_counts[0];
  o->n_candidates = L->last_counts[1];
  o->n_deferred = L->last_counts[2];
  o->window = L->W;
  o->persistent = L->last_persistent;
  o->block = L->last_block;
  o->use_graph = L->last_graph ? 1 : 0;
  o->pipeline_ms = L->last_ms;
  o->window_ms = L->last_win_ms;
  return PL_OK;
}

pl_status pl_last_wide_columns(pl_ctx* c, pl_wide_col* out, int64_t cap, int64_t* n) {
  Lane* L = c ? &c->main : nullptr;
  if (!c || !n) return fail(PL_E_INVAL, "null argument");
  if (!L->tables_valid) return fail(PL_E_INVAL, "no batch on the device");
  const uint32_t nw = std::min(L->last_counts[6], L->wide_cap);
  *n = nw;
  if (!out || !nw) return PL_OK;
  if (cap < (int64_t)nw) return fail(PL_E_INVAL, "%u wide columns, room for %lld", nw, (long long)cap);
  std::vector<WideCol> h(nw);
  HIPCHK(hipMemcpyAsync(h.data(), L->wide.p, sizeof(WideCol) * nw, hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  std::sort(h.begin(), h.end(), [](const WideCol& a, const WideCol& b) { return a.gpos < b.gpos; });
  for (uint32_t k = 0; k < nw; k++) {
    out[k].pos = h[k].gpos;
    for (int b = 0; b < 4; b++) out[k].qual_sum[b] = h[k].qs[b];
  }
  return PL_OK;
}

pl_status pl_last_timing(pl_ctx* c, double* pipeline_ms, double* window_ms) {
  Lane* L = c ? &c->main : nullptr;
  if (!c) return fail(PL_E_INVAL, "null argument");
  if (pipeline_ms) *pipeline_ms = L->last_ms;
  if (window_ms) *window_ms = L->last_win_ms;
  return PL_OK;
}

}  // extern "C"

#include "pl_serve.hip"
#include "pl_vcf.hip"
#include "pl_track.hip"
#include "pl_kmer_kernels.hip"
#include "pl_kmer.hip"
#include "pl_asm.hip"
#include "pl_load.hip"
#include "pl_fix.hip"
