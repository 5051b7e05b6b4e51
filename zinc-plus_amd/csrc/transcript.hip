// transcript.hip -- the Fiat-Shamir transcript and the proof byte stream
// (host code).
//
// Reference: transcript/src/lib.rs:10-72 (Blake3Transcript: challenges are
// the XOF output of the running BLAKE3 state, then absorbed between the tag
// bytes 0x12 and 0x34; get_prime absorbs every candidate),
// transcript/src/traits.rs:358-420 (get_challenges, get_random_field_cfg,
// absorb_slice with tags 0x6/0x7, absorb_random_field with tags
// 0x3 0x5 0x1 0x3), zip-plus/src/pcs_transcript.rs:27-221
// (PcsProverTranscript: squeeze_challenge_idx, write_field_elements,
// write_const_many, write_merkle_proof), primality/src/lib.rs:9-20
// (MillerRabin: the strong probable-prime test to base 2).
//
// The hasher is an incremental BLAKE3 (the published specification: 1 KiB
// chunks, a stack of chunk chaining values merged as chunks complete, the
// root node recompressed with increasing counters for extended output).
#include <immintrin.h>
#include <sys/mman.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "common.h"
#include "field.h"

namespace zpi {
namespace {

const uint32_t kIV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                         0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr uint8_t kSchedule[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum : uint32_t { kChunkStart = 1, kChunkEnd = 2, kParent = 4, kRoot = 8 };

inline uint32_t rotr(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

void compress_scalar(const uint32_t cv[8], const uint32_t m_in[16], uint64_t counter,
                     uint32_t len, uint32_t flags, uint32_t out[16]) {
  uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    kIV[0], kIV[1], kIV[2], kIV[3], (uint32_t)counter, (uint32_t)(counter >> 32),
                    len, flags};
  uint32_t m[16];
  std::memcpy(m, m_in, 64);
  auto g = [&](int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] += v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] += v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] += v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] += v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
  };
  for (int r = 0; r < 7; r++) {
    g(0, 4, 8, 12, m[0], m[1]);
    g(1, 5, 9, 13, m[2], m[3]);
    g(2, 6, 10, 14, m[4], m[5]);
    g(3, 7, 11, 15, m[6], m[7]);
    g(0, 5, 10, 15, m[8], m[9]);
    g(1, 6, 11, 12, m[10], m[11]);
    g(2, 7, 8, 13, m[12], m[13]);
    g(3, 4, 9, 14, m[14], m[15]);
    uint32_t t[16];
    for (int i = 0; i < 16; i++) t[i] = m[kSchedule[i]];
    std::memcpy(m, t, 64);
  }
  for (int i = 0; i < 8; i++) {
    out[i] = v[i] ^ v[i + 8];
    out[i + 8] = v[i + 8] ^ cv[i];
  }
}

// The message word order of each of the seven rounds: round r reads
// m[kMsgOrder.idx[r][i]] where the scalar compression reads its permuted m[i].
struct MsgOrder {
  uint8_t idx[7][16];
};
constexpr MsgOrder msg_order() {
  MsgOrder o{};
  for (int i = 0; i < 16; i++) o.idx[0][i] = (uint8_t)i;
  for (int r = 1; r < 7; r++)
    for (int i = 0; i < 16; i++) o.idx[r][i] = o.idx[r - 1][kSchedule[i]];
  return o;
}
constexpr MsgOrder kMsgOrder = msg_order();

// The same compression with the state as four rows of four words in 128-bit
// registers: the column step runs G on whole rows; the diagonal step first
// rotates rows 1, 2 and 3 left by one, two and three lanes, so that lane j
// holds the j-th diagonal, and rotates them back afterwards.  The 32-bit
// rotates are AVX-512VL's vprord.
__attribute__((target("avx512f,avx512vl"))) void compress_rows(const uint32_t cv[8],
                                                              const uint32_t m[16],
                                                              uint64_t counter, uint32_t len,
                                                              uint32_t flags, uint32_t out[16]) {
  const __m128i c0 = _mm_loadu_si128((const __m128i*)cv);
  const __m128i c1 = _mm_loadu_si128((const __m128i*)(cv + 4));
  __m128i r0 = c0, r1 = c1;
  __m128i r2 = _mm_loadu_si128((const __m128i*)kIV);
  __m128i r3 = _mm_setr_epi32((int)(uint32_t)counter, (int)(uint32_t)(counter >> 32), (int)len,
                              (int)flags);
#define ZP_GROW(x, y)                                  \
  r0 = _mm_add_epi32(_mm_add_epi32(r0, r1), x);        \
  r3 = _mm_ror_epi32(_mm_xor_si128(r3, r0), 16);       \
  r2 = _mm_add_epi32(r2, r3);                          \
  r1 = _mm_ror_epi32(_mm_xor_si128(r1, r2), 12);       \
  r0 = _mm_add_epi32(_mm_add_epi32(r0, r1), y);        \
  r3 = _mm_ror_epi32(_mm_xor_si128(r3, r0), 8);        \
  r2 = _mm_add_epi32(r2, r3);                          \
  r1 = _mm_ror_epi32(_mm_xor_si128(r1, r2), 7);
  for (int r = 0; r < 7; r++) {
    const uint8_t* o = kMsgOrder.idx[r];
    const __m128i x0 = _mm_setr_epi32((int)m[o[0]], (int)m[o[2]], (int)m[o[4]], (int)m[o[6]]);
    const __m128i y0 = _mm_setr_epi32((int)m[o[1]], (int)m[o[3]], (int)m[o[5]], (int)m[o[7]]);
    const __m128i x1 = _mm_setr_epi32((int)m[o[8]], (int)m[o[10]], (int)m[o[12]], (int)m[o[14]]);
    const __m128i y1 = _mm_setr_epi32((int)m[o[9]], (int)m[o[11]], (int)m[o[13]], (int)m[o[15]]);
    ZP_GROW(x0, y0)
    r1 = _mm_shuffle_epi32(r1, _MM_SHUFFLE(0, 3, 2, 1));
    r2 = _mm_shuffle_epi32(r2, _MM_SHUFFLE(1, 0, 3, 2));
    r3 = _mm_shuffle_epi32(r3, _MM_SHUFFLE(2, 1, 0, 3));
    ZP_GROW(x1, y1)
    r1 = _mm_shuffle_epi32(r1, _MM_SHUFFLE(2, 1, 0, 3));
    r2 = _mm_shuffle_epi32(r2, _MM_SHUFFLE(1, 0, 3, 2));
    r3 = _mm_shuffle_epi32(r3, _MM_SHUFFLE(0, 3, 2, 1));
  }
#undef ZP_GROW
  _mm_storeu_si128((__m128i*)out, _mm_xor_si128(r0, r2));
  _mm_storeu_si128((__m128i*)(out + 4), _mm_xor_si128(r1, r3));
  _mm_storeu_si128((__m128i*)(out + 8), _mm_xor_si128(r2, c0));
  _mm_storeu_si128((__m128i*)(out + 12), _mm_xor_si128(r3, c1));
}

// Eight full chunks at once, one per 32-bit lane of AVX2 vectors: chunk i
// at p + 1024 i with chunk counter ctr + i; its chaining value goes to
// cvs[8 i .. 8 i + 8).  The message words of a block are gathered across
// the chunks; rotations by 16 and 8 are byte shuffles.
__attribute__((target("avx2"))) void hash8_chunks_avx2(const uint8_t* p, uint64_t ctr,
                                                      uint32_t* cvs) {
  const __m256i rot16 = _mm256_setr_epi8(2, 3, 0, 1, 6, 7, 4, 5, 10, 11, 8, 9, 14, 15, 12, 13,
                                         2, 3, 0, 1, 6, 7, 4, 5, 10, 11, 8, 9, 14, 15, 12, 13);
  const __m256i rot8 = _mm256_setr_epi8(1, 2, 3, 0, 5, 6, 7, 4, 9, 10, 11, 8, 13, 14, 15, 12,
                                        1, 2, 3, 0, 5, 6, 7, 4, 9, 10, 11, 8, 13, 14, 15, 12);
  const __m256i idx = _mm256_setr_epi32(0, 256, 512, 768, 1024, 1280, 1536, 1792);
  alignas(32) uint32_t lo[8], hi[8];
  for (int i = 0; i < 8; i++) {
    lo[i] = (uint32_t)(ctr + i);
    hi[i] = (uint32_t)((ctr + i) >> 32);
  }
  const __m256i ctr_lo = _mm256_load_si256((const __m256i*)lo);
  const __m256i ctr_hi = _mm256_load_si256((const __m256i*)hi);
  __m256i h[8];
  for (int w = 0; w < 8; w++) h[w] = _mm256_set1_epi32((int)kIV[w]);
#define ZP_ROTR12(x) _mm256_or_si256(_mm256_srli_epi32(x, 12), _mm256_slli_epi32(x, 20))
#define ZP_ROTR7(x) _mm256_or_si256(_mm256_srli_epi32(x, 7), _mm256_slli_epi32(x, 25))
#define ZP_G8(a, bb, c, d, x, y)                                                  \
  v[a] = _mm256_add_epi32(_mm256_add_epi32(v[a], v[bb]), x);                     \
  v[d] = _mm256_shuffle_epi8(_mm256_xor_si256(v[d], v[a]), rot16);               \
  v[c] = _mm256_add_epi32(v[c], v[d]);                                           \
  v[bb] = ZP_ROTR12(_mm256_xor_si256(v[bb], v[c]));                              \
  v[a] = _mm256_add_epi32(_mm256_add_epi32(v[a], v[bb]), y);                     \
  v[d] = _mm256_shuffle_epi8(_mm256_xor_si256(v[d], v[a]), rot8);                \
  v[c] = _mm256_add_epi32(v[c], v[d]);                                           \
  v[bb] = ZP_ROTR7(_mm256_xor_si256(v[bb], v[c]));
  for (int b = 0; b < 16; b++) {
    __m256i m[16];
    for (int w = 0; w < 16; w++)
      m[w] = _mm256_i32gather_epi32((const int*)(p + 64 * b + 4 * w), idx, 4);
    __m256i v[16];
    for (int w = 0; w < 8; w++) v[w] = h[w];
    for (int w = 0; w < 4; w++) v[8 + w] = _mm256_set1_epi32((int)kIV[w]);
    v[12] = ctr_lo;
    v[13] = ctr_hi;
    v[14] = _mm256_set1_epi32(64);
    v[15] = _mm256_set1_epi32((int)((b == 0 ? kChunkStart : 0u) | (b == 15 ? kChunkEnd : 0u)));
    for (int r = 0; r < 7; r++) {
      ZP_G8(0, 4, 8, 12, m[0], m[1])
      ZP_G8(1, 5, 9, 13, m[2], m[3])
      ZP_G8(2, 6, 10, 14, m[4], m[5])
      ZP_G8(3, 7, 11, 15, m[6], m[7])
      ZP_G8(0, 5, 10, 15, m[8], m[9])
      ZP_G8(1, 6, 11, 12, m[10], m[11])
      ZP_G8(2, 7, 8, 13, m[12], m[13])
      ZP_G8(3, 4, 9, 14, m[14], m[15])
      __m256i t[16];
      for (int i = 0; i < 16; i++) t[i] = m[kSchedule[i]];
      for (int i = 0; i < 16; i++) m[i] = t[i];
    }
    for (int w = 0; w < 8; w++) h[w] = _mm256_xor_si256(v[w], v[w + 8]);
  }
  alignas(32) uint32_t tmp[8][8];
  for (int w = 0; w < 8; w++) _mm256_store_si256((__m256i*)tmp[w], h[w]);
  for (int i = 0; i < 8; i++)
    for (int w = 0; w < 8; w++) cvs[8 * i + w] = tmp[w][i];
#undef ZP_G8
#undef ZP_ROTR7
#undef ZP_ROTR12
}

// Sixteen full chunks at once with AVX-512 (native 32-bit rotates); the
// same layout as hash8_chunks_avx2.
__attribute__((target("avx512f"))) void hash16_chunks_avx512(const uint8_t* p, uint64_t ctr,
                                                            uint32_t* cvs) {
  const __m512i idx = _mm512_setr_epi32(0, 256, 512, 768, 1024, 1280, 1536, 1792, 2048, 2304,
                                        2560, 2816, 3072, 3328, 3584, 3840);
  alignas(64) uint32_t lo[16], hi[16];
  for (int i = 0; i < 16; i++) {
    lo[i] = (uint32_t)(ctr + i);
    hi[i] = (uint32_t)((ctr + i) >> 32);
  }
  const __m512i ctr_lo = _mm512_load_si512(lo);
  const __m512i ctr_hi = _mm512_load_si512(hi);
  __m512i h[8];
  for (int w = 0; w < 8; w++) h[w] = _mm512_set1_epi32((int)kIV[w]);
#define ZP_G16(a, bb, c, d, x, y)                                               \
  v[a] = _mm512_add_epi32(_mm512_add_epi32(v[a], v[bb]), x);                   \
  v[d] = _mm512_ror_epi32(_mm512_xor_si512(v[d], v[a]), 16);                   \
  v[c] = _mm512_add_epi32(v[c], v[d]);                                         \
  v[bb] = _mm512_ror_epi32(_mm512_xor_si512(v[bb], v[c]), 12);                 \
  v[a] = _mm512_add_epi32(_mm512_add_epi32(v[a], v[bb]), y);                   \
  v[d] = _mm512_ror_epi32(_mm512_xor_si512(v[d], v[a]), 8);                    \
  v[c] = _mm512_add_epi32(v[c], v[d]);                                         \
  v[bb] = _mm512_ror_epi32(_mm512_xor_si512(v[bb], v[c]), 7);
  for (int b = 0; b < 16; b++) {
    __m512i m[16];
    for (int w = 0; w < 16; w++)
      m[w] = _mm512_i32gather_epi32(idx, (const void*)(p + 64 * b + 4 * w), 4);
    __m512i v[16];
    for (int w = 0; w < 8; w++) v[w] = h[w];
    for (int w = 0; w < 4; w++) v[8 + w] = _mm512_set1_epi32((int)kIV[w]);
    v[12] = ctr_lo;
    v[13] = ctr_hi;
    v[14] = _mm512_set1_epi32(64);
    v[15] = _mm512_set1_epi32((int)((b == 0 ? kChunkStart : 0u) | (b == 15 ? kChunkEnd : 0u)));
    for (int r = 0; r < 7; r++) {
      ZP_G16(0, 4, 8, 12, m[0], m[1])
      ZP_G16(1, 5, 9, 13, m[2], m[3])
      ZP_G16(2, 6, 10, 14, m[4], m[5])
      ZP_G16(3, 7, 11, 15, m[6], m[7])
      ZP_G16(0, 5, 10, 15, m[8], m[9])
      ZP_G16(1, 6, 11, 12, m[10], m[11])
      ZP_G16(2, 7, 8, 13, m[12], m[13])
      ZP_G16(3, 4, 9, 14, m[14], m[15])
      __m512i t[16];
      for (int i = 0; i < 16; i++) t[i] = m[kSchedule[i]];
      for (int i = 0; i < 16; i++) m[i] = t[i];
    }
    for (int w = 0; w < 8; w++) h[w] = _mm512_xor_si512(v[w], v[w + 8]);
  }
  alignas(64) uint32_t tmp[8][16];
  for (int w = 0; w < 8; w++) _mm512_store_si512(tmp[w], h[w]);
  for (int i = 0; i < 16; i++)
    for (int w = 0; w < 8; w++) cvs[8 * i + w] = tmp[w][i];
#undef ZP_G16
}

// ZIPGPU_HOST_SIMD=0 in the environment keeps the hasher on its scalar code
// (what a host without AVX2 runs); the digests are the same either way.
bool host_simd() {
  static const bool yes = [] {
    const char* e = std::getenv("ZIPGPU_HOST_SIMD");
    return !(e && e[0] == '0');
  }();
  return yes;
}

bool host_has_avx2() {
  static const bool yes = host_simd() && __builtin_cpu_supports("avx2");
  return yes;
}

bool host_has_avx512() {
  static const bool yes = host_simd() && __builtin_cpu_supports("avx512f");
  return yes;
}

bool host_has_avx512vl() {
  static const bool yes = host_has_avx512() && __builtin_cpu_supports("avx512vl");
  return yes;
}

void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t len,
              uint32_t flags, uint32_t out[16]) {
  if (host_has_avx512vl()) compress_rows(cv, m, counter, len, flags, out);
  else compress_scalar(cv, m, counter, len, flags, out);
}

void words_le(const uint8_t* b, uint32_t* w, int n) {
  for (int i = 0; i < n; i++)
    w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
           ((uint32_t)b[4 * i + 3] << 24);
}

// The node whose compression gives a chaining value or the root output.
struct Node {
  uint32_t cv[8];
  uint32_t block[16];
  uint64_t counter;
  uint32_t len, flags;
  void chaining(uint32_t out[8]) const {
    uint32_t o[16];
    compress(cv, block, counter, len, flags, o);
    std::memcpy(out, o, 32);
  }
};

Node parent_node(const uint32_t l[8], const uint32_t r[8]) {
  Node n;
  std::memcpy(n.cv, kIV, 32);
  std::memcpy(n.block, l, 32);
  std::memcpy(n.block + 8, r, 32);
  n.counter = 0;
  n.len = 64;
  n.flags = kParent;
  return n;
}

}  // namespace

class Blake3 {
 public:
  Blake3() { reset_chunk(0); }

  void update(const uint8_t* p, size_t n) {
    while (n) {
      if (chunk_len() == 1024) {
        uint32_t cv[8];
        chunk_node().chaining(cv);
        push_chunk(cv, chunk_ + 1);
        reset_chunk(chunk_ + 1);
      }
      // At a chunk boundary with more than sixteen (eight) chunks of input
      // left, so none of them is the last chunk: sixteen (eight) at once.
      while (chunk_len() == 0 && n > 16 * 1024 && host_has_avx512()) {
        uint32_t cvs[128];
        hash16_chunks_avx512(p, chunk_, cvs);
        for (int i = 0; i < 16; i++) push_chunk_cv(cvs + 8 * i);
        p += 16 * 1024;
        n -= 16 * 1024;
      }
      while (chunk_len() == 0 && n > 8 * 1024 && host_has_avx2()) {
        uint32_t cvs[64];
        hash8_chunks_avx2(p, chunk_, cvs);
        for (int i = 0; i < 8; i++) push_chunk_cv(cvs + 8 * i);
        p += 8 * 1024;
        n -= 8 * 1024;
      }
      if (buf_len_ == 64) {  // a full block is compressed only once more input follows
        uint32_t m[16], o[16];
        words_le(buf_, m, 16);
        compress(cv_, m, chunk_, 64, blocks_ == 0 ? kChunkStart : 0, o);
        std::memcpy(cv_, o, 32);
        blocks_++;
        buf_len_ = 0;
      }
      const size_t take = std::min<size_t>(n, 64 - buf_len_);
      std::memcpy(buf_ + buf_len_, p, take);
      buf_len_ += (uint32_t)take;
      p += take;
      n -= take;
    }
  }

  // finalize_xof().fill(out): does not change the state.
  void xof(uint8_t* out, size_t n) const {
    Node node = chunk_node();
    for (size_t i = stack_.size(); i-- > 0;) {
      uint32_t cv[8];
      node.chaining(cv);
      node = parent_node(stack_[i].w, cv);
    }
    for (uint64_t blk = 0; n; blk++) {
      uint32_t o[16];
      compress(node.cv, node.block, blk, node.len, node.flags | kRoot, o);
      uint8_t bytes[64];
      for (int i = 0; i < 16; i++)
        for (int k = 0; k < 4; k++) bytes[4 * i + k] = (uint8_t)(o[i] >> (8 * k));
      const size_t take = std::min<size_t>(n, 64);
      std::memcpy(out, bytes, take);
      out += take;
      n -= take;
    }
  }

  // Bytes absorbed so far.
  uint64_t total_len() const { return chunk_ * 1024 + chunk_len(); }
  size_t chunk_len() const { return (size_t)blocks_ * 64 + buf_len_; }
  uint64_t chunk_index() const { return chunk_; }

  // With the current chunk full and more input to follow: compress it and
  // push its chaining value, leaving an empty chunk (what update does at the
  // next byte).
  void flush_full_chunk() {
    uint32_t cv[8];
    chunk_node().chaining(cv);
    push_chunk(cv, chunk_ + 1);
    reset_chunk(chunk_ + 1);
  }

  // With an empty current chunk and more input to follow: take the chaining
  // value of the full chunk at this chunk index, computed elsewhere.
  void push_chunk_cv(const uint32_t cv[8]) {
    push_chunk(cv, chunk_ + 1);
    reset_chunk(chunk_ + 1);
  }

 private:
  struct Cv {
    uint32_t w[8];
  };
  uint32_t cv_[8];
  uint64_t chunk_ = 0;
  uint8_t buf_[64];
  uint32_t buf_len_ = 0, blocks_ = 0;
  std::vector<Cv> stack_;

  void reset_chunk(uint64_t idx) {
    std::memcpy(cv_, kIV, 32);
    chunk_ = idx;
    buf_len_ = 0;
    blocks_ = 0;
  }

  Node chunk_node() const {
    Node n;
    std::memcpy(n.cv, cv_, 32);
    uint8_t b[64] = {0};
    std::memcpy(b, buf_, buf_len_);
    words_le(b, n.block, 16);
    n.counter = chunk_;
    n.len = buf_len_;
    n.flags = (blocks_ == 0 ? kChunkStart : 0) | kChunkEnd;
    return n;
  }

  // Merge completed subtrees: one per trailing zero bit of the chunk count.
  void push_chunk(const uint32_t cv[8], uint64_t total) {
    Cv c;
    std::memcpy(c.w, cv, 32);
    while ((total & 1) == 0) {
      Node n = parent_node(stack_.back().w, c.w);
      stack_.pop_back();
      n.chaining(c.w);
      total >>= 1;
    }
    stack_.push_back(c);
  }
};

// SPRP test to base 2 of an odd n > 3 (primality/src/lib.rs:13-19).
bool sprp2(const uint64_t n[4]) {
  if ((n[0] & 1) == 0) return false;
  if ((n[1] | n[2] | n[3]) == 0 && n[0] < 5) return n[0] == 3;
  const FpParams P = fp_params(n);
  uint64_t d[4] = {n[0] - 1, n[1], n[2], n[3]};  // n odd: no borrow
  int s = 0;
  while ((d[0] & 1) == 0) {
    for (int i = 0; i < 3; i++) d[i] = (d[i] >> 1) | (d[i + 1] << 63);
    d[3] >>= 1;
    s++;
  }
  const Fp one = fp_one(P);
  const Fp minus_one = fp_neg(P, one);
  Fp x = one;  // 2^d, left to right: square, then double for a set bit
  int top = 255;
  while (top >= 0 && !((d[top / 64] >> (top % 64)) & 1)) top--;
  for (int b = top; b >= 0; b--) {
    x = fp_mul(P, x, x);
    if ((d[b / 64] >> (b % 64)) & 1) x = fp_add(P, x, x);
  }
  auto eq = [](const Fp& a, const Fp& b) { return std::memcmp(a.w, b.w, 32) == 0; };
  if (eq(x, one) || eq(x, minus_one)) return true;
  for (int r = 1; r < s; r++) {
    x = fp_mul(P, x, x);
    if (eq(x, minus_one)) return true;
    if (eq(x, one)) return false;
  }
  return false;
}

}  // namespace zpi

namespace zpi {

// Copies n bytes; large copies are split over host threads, which also
// spreads the first-touch page faults of a freshly grown destination.
void par_memcpy(void* dst, const void* src, size_t n) {
  const size_t kMin = 8u << 20;
  if (n < kMin) {
    std::memcpy(dst, src, n);
    return;
  }
  static const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (hw == 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const unsigned T = (unsigned)std::min<size_t>(hw, n / (2u << 20));
  const size_t per = (n / T + 4095) & ~(size_t)4095;
  std::vector<std::thread> th;
  for (unsigned i = 1; i < T; i++) {
    const size_t off = i * per;
    if (off >= n) break;
    th.emplace_back([=] {
      std::memcpy((uint8_t*)dst + off, (const uint8_t*)src + off, std::min(per, n - off));
    });
  }
  std::memcpy(dst, src, std::min(per, n));
  for (auto& x : th) x.join();
}

// Large stream storage: pinned host buffers (hipHostMalloc), kept for reuse
// when a stream lets go of them.  A proof with column openings reaches tens
// of MiB; fresh pages cost a fault and a kernel zeroing each on first touch
// (a 77 MB stream: ~19,000 faults with 4 KiB pages), and pinned storage
// lets the prover copy the proof's device parts straight into the stream.
// At most kCacheBytes stay cached; the rest is released.
class PinnedStore {
 public:
  // A buffer of at least `want` bytes (rounded to 2 MiB), or null when
  // pinned memory is not available.
  static uint8_t* acquire(size_t want, size_t* cap) {
    const size_t c = (want + kGranule - 1) & ~(kGranule - 1);
    {
      std::lock_guard<std::mutex> g(mu());
      auto& fl = free_list();
      size_t best = fl.size();
      for (size_t i = 0; i < fl.size(); i++)
        if (fl[i].second >= c && (best == fl.size() || fl[i].second < fl[best].second)) best = i;
      if (best < fl.size()) {
        const auto e = fl[best];
        fl.erase(fl.begin() + (long)best);
        cached() -= e.second;
        *cap = e.second;
        return (uint8_t*)e.first;
      }
    }
    void* q = nullptr;
    if (hipHostMalloc(&q, c, hipHostMallocDefault) != hipSuccess || !q) {
      (void)hipGetLastError();
      return nullptr;
    }
    *cap = c;
    return (uint8_t*)q;
  }
  static void release(uint8_t* p, size_t cap) {
    if (!p) return;
    {
      std::lock_guard<std::mutex> g(mu());
      if (cached() + cap <= kCacheBytes) {
        free_list().push_back({p, cap});
        cached() += cap;
        return;
      }
    }
    (void)hipHostFree(p);
  }

 private:
  static constexpr size_t kGranule = 2u << 20, kCacheBytes = 1ull << 30;
  // Never destroyed: a stream released during process exit still finds them.
  static std::mutex& mu() {
    static std::mutex* m = new std::mutex;
    return *m;
  }
  static std::vector<std::pair<void*, size_t>>& free_list() {
    static auto* v = new std::vector<std::pair<void*, size_t>>;
    return *v;
  }
  static size_t& cached() {
    static size_t b = 0;
    return b;
  }
};

// The proof stream's bytes.  Small streams grow with realloc; from 4 MiB on
// the storage comes from PinnedStore (or, without pinned memory, 2 MiB-
// aligned pageable memory advised for transparent huge pages).
struct ByteBuf {
  static constexpr size_t kHuge = 2u << 20, kBig = 4u << 20;
  uint8_t* p = nullptr;
  size_t n = 0, cap = 0;
  bool pinned = false;  // p came from PinnedStore
  ByteBuf() = default;
  ByteBuf(const ByteBuf& o) {
    reserve(o.n);
    if (o.n) par_memcpy(p, o.p, o.n);
    n = o.n;
  }
  ByteBuf& operator=(const ByteBuf&) = delete;
  ~ByteBuf() { drop(p, cap, pinned); }
  static void drop(uint8_t* q, size_t c, bool pin) {
    if (pin) PinnedStore::release(q, c);
    else std::free(q);
  }
  void reserve(size_t want) {
    if (want <= cap) return;
    size_t c = std::max(want, cap * 2);
    if (c < kBig) {
      uint8_t* q = (uint8_t*)std::realloc(p, std::max<size_t>(c, 64));
      if (!q) throw std::bad_alloc();
      p = q;
      cap = std::max<size_t>(c, 64);
      return;
    }
    size_t got = 0;
    uint8_t* q = PinnedStore::acquire(c, &got);
    bool pin = q != nullptr;
    if (!q) {
      got = (c + kHuge - 1) & ~(kHuge - 1);
      void* v = nullptr;
      if (posix_memalign(&v, kHuge, got) != 0) throw std::bad_alloc();
      (void)madvise(v, got, MADV_HUGEPAGE);
      q = (uint8_t*)v;
    }
    if (n) par_memcpy(q, p, n);
    drop(p, cap, pinned);
    p = q;
    cap = got;
    pinned = pin;
  }
  // Appends k uninitialised bytes and returns where they start.
  uint8_t* grow(size_t k) {
    reserve(n + k);
    uint8_t* at = p + n;
    n += k;
    return at;
  }
};

}  // namespace zpi

struct zp_transcript {
  zpi::Blake3 fs;
  zpi::ByteBuf stream;

  void absorb(const void* p, size_t n) { fs.update((const uint8_t*)p, n); }
  void absorb_tag(uint8_t t) { fs.update(&t, 1); }
  void challenge(uint8_t* out, size_t n) {
    fs.xof(out, n);
    if (n <= 62) {  // 0x12, the bytes, 0x34 in one update
      uint8_t rec[64];
      rec[0] = 0x12;
      std::memcpy(rec + 1, out, n);
      rec[n + 1] = 0x34;
      absorb(rec, n + 2);
      return;
    }
    absorb_tag(0x12);
    absorb(out, n);
    absorb_tag(0x34);
  }
  void put(const void* p, size_t n) {
    if (n) zpi::par_memcpy(stream.grow(n), p, n);
  }
};

namespace zpi {

void ts_challenge(zp_transcript* t, uint8_t* out, size_t n) { t->challenge(out, n); }

uint64_t ts_squeeze_idx(zp_transcript* t, uint64_t cap) {
  uint8_t b[4];
  t->challenge(b, 4);
  const uint32_t v = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) |
                     ((uint32_t)b[3] << 24);
  return v % cap;
}

// write_field_elements (pcs_transcript.rs:99-142): Uint<4> has no length
// prefix; each element is absorbed (absorb_random_field) and written as
// modulus bytes, then inner bytes.
// The absorbed records (68 bytes each) and the stream bytes (64 each) are
// laid out in host buffers first and handed over in one call each: BLAKE3
// is a streaming hash, so the digest does not depend on how the input is
// split.
void ts_write_field_elements(zp_transcript* t, const uint64_t mod[4], const Fp* v, size_t n) {
  constexpr size_t kRec = 68, kOut = 64, kBatch = 1024;
  std::vector<uint8_t> rec(kRec * std::min(n, kBatch));
  uint8_t* out = n ? t->stream.grow(kOut * n) : nullptr;
  for (size_t i0 = 0; i0 < n; i0 += kBatch) {
    const size_t m = std::min(kBatch, n - i0);
    for (size_t i = 0; i < m; i++) {
      uint8_t* r = rec.data() + kRec * i;
      r[0] = 0x3;
      std::memcpy(r + 1, mod, 32);
      r[33] = 0x5;
      r[34] = 0x1;
      std::memcpy(r + 35, v[i0 + i].w, 32);
      r[67] = 0x3;
      std::memcpy(out + kOut * (i0 + i), mod, 32);
      std::memcpy(out + kOut * (i0 + i) + 32, v[i0 + i].w, 32);
    }
    t->absorb(rec.data(), kRec * m);
  }
}

void ts_put(zp_transcript* t, const void* p, size_t n) { t->put(p, n); }

void ts_reserve(zp_transcript* t, size_t more) { t->stream.reserve(t->stream.n + more); }

// `more` bytes appended to the stream, uninitialised, for a device copy to
// fill: their address if the storage is pinned (else null and nothing is
// appended).  ts_unput(t, more) takes them back after a failed copy.
uint8_t* ts_grow_pinned(zp_transcript* t, size_t more) {
  t->stream.reserve(t->stream.n + more);
  if (!t->stream.pinned) return nullptr;
  return t->stream.grow(more);
}
void ts_unput(zp_transcript* t, size_t n) { t->stream.n -= std::min(n, t->stream.n); }

}  // namespace zpi

using namespace zpi;

extern "C" {

int zp_transcript_create(zp_transcript** out) {
  if (!out) return fail(ZP_ERR_INVALID_PARAM, "null output");
  *out = new zp_transcript();
  return ZP_OK;
}

int zp_transcript_clone(const zp_transcript* t, zp_transcript** out) {
  if (!t || !out) return fail(ZP_ERR_INVALID_PARAM, "null transcript");
  try {
    *out = new zp_transcript(*t);
  } catch (const std::bad_alloc&) {
    return fail(ZP_ERR_OOM, "host allocation failed");
  }
  return ZP_OK;
}

void zp_transcript_destroy(zp_transcript* t) { delete t; }

int zp_transcript_absorb_slice(zp_transcript* t, const uint8_t* bytes, uint64_t len) {
  if (!t || (len && !bytes)) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  t->absorb_tag(0x6);
  t->absorb(bytes, len);
  t->absorb_tag(0x7);
  return ZP_OK;
}

int zp_transcript_absorb_entries(zp_transcript* t, const uint8_t* bytes, uint64_t entry_bytes,
                                 uint64_t count) {
  if (!t || (count && entry_bytes && !bytes)) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  // The framed entries (0x6 || entry || 0x7) are assembled in a buffer and
  // absorbed in large pieces, so the hasher can take many chunks at once.
  constexpr size_t kBuf = 256u << 10;
  const uint64_t F = entry_bytes + 2;
  if (F > kBuf) {
    for (uint64_t i = 0; i < count; i++) {
      t->absorb_tag(0x6);
      t->absorb(bytes + i * entry_bytes, entry_bytes);
      t->absorb_tag(0x7);
    }
    return ZP_OK;
  }
  std::vector<uint8_t> buf;
  try {
    buf.resize(kBuf);
  } catch (const std::bad_alloc&) {
    return fail(ZP_ERR_OOM, "host allocation failed");
  }
  const uint64_t per = kBuf / F;
  for (uint64_t i0 = 0; i0 < count; i0 += per) {
    const uint64_t k = std::min(per, count - i0);
    uint8_t* q = buf.data();
    for (uint64_t i = 0; i < k; i++) {
      *q++ = 0x6;
      if (entry_bytes) std::memcpy(q, bytes + (i0 + i) * entry_bytes, entry_bytes);
      q += entry_bytes;
      *q++ = 0x7;
    }
    t->absorb(buf.data(), (size_t)(q - buf.data()));
  }
  return ZP_OK;
}

namespace {

// Absorbs bytes [qb, qe) of the framed stream of `count` entries of E bytes
// (0x6 || entry || 0x7 each) whose data lies in device memory.
int feed_framed(zp_transcript* t, const uint8_t* d, uint64_t E, uint64_t qb, uint64_t qe,
                hipStream_t s) {
  if (qe <= qb) return ZP_OK;
  const uint64_t F = E + 2;
  auto data_pos = [&](uint64_t q) {  // data bytes before stream offset q
    const uint64_t i = q / F, u = q % F;
    return i * E + (u == 0 ? 0 : std::min(u - 1, E));
  };
  const uint64_t lo = data_pos(qb), hi = data_pos(qe);
  std::vector<uint8_t> hd(hi - lo), out(qe - qb);
  if (hi > lo) {
    ZP_HIP(hipMemcpyAsync(hd.data(), d + lo, hi - lo, hipMemcpyDeviceToHost, s));
    ZP_HIP(hipStreamSynchronize(s));
  }
  uint64_t i = qb / F, u = qb % F;
  for (uint64_t q = qb; q < qe; q++) {
    out[q - qb] = u == 0 ? 0x6 : u == E + 1 ? 0x7 : hd[i * E + u - 1 - lo];
    if (++u == F) {
      u = 0;
      i++;
    }
  }
  t->absorb(out.data(), out.size());
  return ZP_OK;
}

// Below this many stream bytes the device absorb copies the data and hashes
// it on the host.
constexpr uint64_t kDeviceAbsorbMin = 32 * 1024;

}  // namespace

int zp_transcript_absorb_entries_device(zp_transcript* t, const void* d_bytes,
                                        uint64_t entry_bytes, uint64_t count, zp_stream zs) {
  const hipStream_t stream = (hipStream_t)zs;
  if (!t || (count && entry_bytes && !d_bytes)) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  const uint64_t E = entry_bytes, F = E + 2;
  if (F < E || (count && F > UINT64_MAX / count))
    return fail(ZP_ERR_INVALID_PARAM, "stream too long");
  const uint64_t T = count * F;
  const uint8_t* d = (const uint8_t*)d_bytes;
  if (T < kDeviceAbsorbMin) return feed_framed(t, d, E, 0, T, stream);
  // Fill the current chunk on the host, then hash every further full chunk
  // that is not the stream's last on the device; the host keeps the tail.
  const uint64_t h = t->fs.total_len() == 0 ? 0 : 1024 - t->fs.chunk_len();
  ZP_TRY(feed_framed(t, d, E, 0, h, stream));
  if (t->fs.total_len() != 0) t->fs.flush_full_chunk();
  const uint64_t n = (T - h - 1) / 1024;
  std::vector<uint32_t> cvs(n * 8);
  uint32_t* d_cvs = nullptr;
  ZP_TRY(dev_alloc((void**)&d_cvs, n * 32, stream));
  int rc = launch_framed_chunk_cvs(d, E, h, t->fs.chunk_index(), n, d_cvs, stream);
  if (rc == ZP_OK && (hipMemcpyAsync(cvs.data(), d_cvs, n * 32, hipMemcpyDeviceToHost, stream) !=
                          hipSuccess ||
                      hipStreamSynchronize(stream) != hipSuccess))
    rc = fail(ZP_ERR_DEVICE, "copy of chunk chaining values failed");
  dev_free(d_cvs, stream);
  if (rc != ZP_OK) return rc;
  for (uint64_t k = 0; k < n; k++) t->fs.push_chunk_cv(&cvs[k * 8]);
  return feed_framed(t, d, E, h + 1024 * n, T, stream);
}

int zp_transcript_challenge(zp_transcript* t, uint8_t* h_out, uint32_t num_bytes) {
  if (!t || (num_bytes && !h_out)) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  t->challenge(h_out, num_bytes);
  return ZP_OK;
}

int zp_transcript_squeeze_idx(zp_transcript* t, uint64_t cap, uint64_t* idx) {
  if (!t || !idx) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  if (cap == 0) return fail(ZP_ERR_SHAPE, "Challenge cap is zero");
  *idx = ts_squeeze_idx(t, cap);
  return ZP_OK;
}

int zp_transcript_random_field_cfg(zp_transcript* t, uint8_t h_modulus[32]) {
  if (!t || !h_modulus) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  for (;;) {
    uint8_t b[32];
    t->fs.xof(b, 32);  // gen_random: fill, absorb without tags
    t->absorb(b, 32);
    uint64_t c[4];
    std::memcpy(c, b, 32);
    if ((c[0] | c[1] | c[2] | c[3]) == 0) continue;
    if ((c[0] & 1) == 0) {  // even: subtract one across the limbs
      const uint64_t one[4] = {1, 0, 0, 0};
      fp_sub_raw(c, one);
    }
    if (sprp2(c)) {
      std::memcpy(h_modulus, c, 32);
      return ZP_OK;
    }
  }
}

int zp_transcript_write(zp_transcript* t, const uint8_t* bytes, uint64_t len) {
  if (!t || (len && !bytes)) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  t->put(bytes, len);
  return ZP_OK;
}

uint64_t zp_transcript_stream_len(const zp_transcript* t) { return t ? t->stream.n : 0; }

int zp_transcript_stream(const zp_transcript* t, uint8_t* h_out, uint64_t cap) {
  if (!t || (cap && !h_out)) return fail(ZP_ERR_INVALID_PARAM, "null argument");
  if (cap < t->stream.n) return fail(ZP_ERR_SHAPE, "buffer smaller than the stream");
  if (t->stream.n) par_memcpy(h_out, t->stream.p, t->stream.n);
  return ZP_OK;
}

}  // extern "C"
