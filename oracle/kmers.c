/*
 * kmers.c — restatement of dada2-rs src/kmers.rs (TEST INFRASTRUCTURE ONLY).
 *
 * Encoding A=1 C=2 G=3 T=4; k-mer index = sum (nt-1) * 4^(k-1-i)
 * (kmers.rs:9-11, encode_kmer kmers.rs:55-66).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* encode_kmer, kmers.rs:55-66: -1 when any base is outside 1..=4 */
static long encode_kmer(const uint8_t* w, int k) {
  long kmer = 0;
  for (int i = 0; i < k; ++i) {
    unsigned nt = w[i];
    if (nt < 1 || nt > 4) return -1;
    kmer = 4 * kmer + (long)(nt - 1);
  }
  return kmer;
}

/* assign_kmer, kmers.rs:79-89: u16 histogram, saturating. */
void orc_assign_kmer(const uint8_t* seq, uint32_t len, int k, uint16_t* out) {
  size_t n = (size_t)1 << (2 * k);
  memset(out, 0, n * sizeof(uint16_t));
  if (len < (uint32_t)k) return;
  for (uint32_t s = 0; s + (uint32_t)k <= len; ++s) {
    long idx = encode_kmer(seq + s, k);
    if (idx >= 0 && out[idx] != 0xFFFF) out[idx] += 1;
  }
}

/* assign_kmer8, kmers.rs:96-101: u16 histogram then min(255). */
void orc_assign_kmer8(const uint8_t* seq, uint32_t len, int k, uint8_t* out) {
  size_t n = (size_t)1 << (2 * k);
  uint16_t* tmp = (uint16_t*)malloc(n * sizeof(uint16_t));
  orc_assign_kmer(seq, len, k, tmp);
  for (size_t i = 0; i < n; ++i) out[i] = tmp[i] > 255 ? 255 : (uint8_t)tmp[i];
  free(tmp);
}

/* assign_kmer_order, kmers.rs:129-135: index per window, 0 if invalid. */
void orc_assign_kmer_order(const uint8_t* seq, uint32_t len, int k, uint16_t* out) {
  if (len < (uint32_t)k) return;
  for (uint32_t s = 0; s + (uint32_t)k <= len; ++s) {
    long idx = encode_kmer(seq + s, k);
    out[s] = idx < 0 ? 0 : (uint16_t)idx;
  }
}

/* kmer_dist, kmers.rs:233-241 */
double orc_kmer_dist(const uint16_t* kv1, uint32_t len1, const uint16_t* kv2, uint32_t len2,
                     int k) {
  size_t n = (size_t)1 << (2 * k);
  uint32_t dot = 0;
  for (size_t i = 0; i < n; ++i) dot += kv1[i] < kv2[i] ? kv1[i] : kv2[i];
  uint32_t mn = len1 < len2 ? len1 : len2;
  double scale = (double)(mn - (uint32_t)k + 1);
  return 1.0 - (double)dot / scale;
}

/* kmer_dist8, kmers.rs:254-269: -1.0 if any elementwise min is 255. */
double orc_kmer_dist8(const uint8_t* kv1, uint32_t len1, const uint8_t* kv2, uint32_t len2,
                      int k) {
  size_t n = (size_t)1 << (2 * k);
  uint32_t dot = 0;
  int overflow = 0;
  for (size_t i = 0; i < n; ++i) {
    uint8_t m = kv1[i] < kv2[i] ? kv1[i] : kv2[i];
    if (m == 255) overflow = 1;
    dot += m;
  }
  if (overflow) return -1.0;
  uint32_t mn = len1 < len2 ? len1 : len2;
  double scale = (double)(mn - (uint32_t)k + 1);
  return 1.0 - (double)dot / scale;
}

/* kmer_dist8_sparse, kmers.rs:280-308: merge-join over ascending indices. */
double orc_kmer_dist8_sparse(const uint16_t* i1, const uint8_t* c1, uint32_t n1, uint32_t len1,
                             const uint16_t* i2, const uint8_t* c2, uint32_t n2, uint32_t len2,
                             int k) {
  uint32_t dot = 0, a = 0, b = 0;
  while (a < n1 && b < n2) {
    if (i1[a] < i2[b]) {
      ++a;
    } else if (i1[a] > i2[b]) {
      ++b;
    } else {
      uint8_t m = c1[a] < c2[b] ? c1[a] : c2[b];
      if (m == 255) return -1.0;
      dot += m;
      ++a;
      ++b;
    }
  }
  uint32_t mn = len1 < len2 ? len1 : len2;
  double scale = (double)(mn - (uint32_t)k + 1);
  return 1.0 - (double)dot / scale;
}

/* kord_dist, kmers.rs:314-328 */
double orc_kord_dist(const uint16_t* kord1, uint32_t len1, const uint16_t* kord2, uint32_t len2,
                     int k) {
  if (len1 != len2) return -1.0;
  uint32_t km1 = k > 0 ? (uint32_t)(k - 1) : 0; /* k.saturating_sub(1) */
  if (len1 < km1) return -1.0;
  uint32_t klen = len1 - km1;
  if (klen == 0) return -1.0;
  uint32_t dot = 0;
  for (uint32_t i = 0; i < klen; ++i) dot += kord1[i] == kord2[i];
  return 1.0 - (double)dot / (double)klen;
}
