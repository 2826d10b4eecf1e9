/* ORACLE — test infrastructure only.  CPU restatement of the R1CS wire
 * format: postcard (the `postcard` 1.x crate, Cargo.toml:124; not vendored)
 * encodings of
 *   R1CS          { num_public_inputs: usize, interner: Interner,
 *                   a, b, c: SparseMatrix }      provekit/common/src/r1cs.rs:59-66
 *   Interner      { values: Vec<Fr> } via serde_ark: one postcard byte string
 *                 holding ark's compressed Vec<Fr> (u64 LE count, then 32-byte
 *                 LE canonical integers)  interner.rs:6-10, utils/serde_ark.rs:11-30
 *   SparseMatrix  (num_rows: usize, num_cols: usize, new_row_indices: Vec<u32>,
 *                 col_deltas: Vec<u32>, values: Vec<InternedFieldElement(usize)>)
 *                 with per-row delta-coded columns   sparse_matrix.rs:73-160
 * Postcard's published format: unsigned integers wider than u8 are LEB128
 * varints (7 bits per byte, low group first, high bit = more follows; at most
 * 5 bytes for u32 and 10 for u64/usize, the last byte holding no bits beyond
 * the type); a sequence is its varint length then its elements; a struct is
 * its fields in order; a newtype is its field; bytes are a varint length then
 * the raw bytes.
 *
 * Matrices are hydrated on the way out: each interned id becomes its Fr
 * (Montgomery form).  Checks beyond the reference's deserializer, which the
 * device loader shares: num_rows equals the row-start count; row starts begin
 * at 0, never decrease and stay within the value count; the delta count
 * equals the value count; columns are below num_cols; ids are below the
 * interner size.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

enum { W_OK = 0, W_BAD = -1, W_SHORT = -2 };

typedef struct {
  const uint8_t* b;
  size_t len, pos;
} rd;

/* LEB128 varint of at most max_bytes bytes whose value fits `bits` bits */
static int rd_varint(rd* r, int max_bytes, int bits, uint64_t* out) {
  uint64_t v = 0;
  for (int i = 0; i < max_bytes; i++) {
    if (r->pos >= r->len) return W_BAD;
    const uint8_t c = r->b[r->pos++];
    const int sh = 7 * i;
    if (i == max_bytes - 1) {
      /* last allowed byte: no continuation, no bits past the type */
      const int room = bits - sh;
      if ((c & 0x80) || (room < 7 && (c >> room) != 0)) return W_BAD;
    }
    v |= (uint64_t)(c & 0x7f) << sh;
    if (!(c & 0x80)) {
      *out = v;
      return W_OK;
    }
  }
  return W_BAD;
}
static int rd_usize(rd* r, uint64_t* out) { return rd_varint(r, 10, 64, out); }
static int rd_u32(rd* r, uint32_t* out) {
  uint64_t v;
  int e = rd_varint(r, 5, 32, &v);
  *out = (uint32_t)v;
  return e;
}

/* Interner (interner.rs:6-10 via serde_ark): canonical -> Montgomery.
 * out may be NULL (count only). */
static int rd_interner(rd* r, ofr* out, size_t cap, size_t* count) {
  uint64_t nbytes, n = 0;
  if (rd_usize(r, &nbytes)) return W_BAD;
  if (nbytes > r->len - r->pos || nbytes < 8) return W_BAD;
  const uint8_t* p = r->b + r->pos;
  for (int i = 0; i < 8; i++) n |= (uint64_t)p[i] << (8 * i);
  if (n > (nbytes - 8) / 32 || nbytes != 8 + 32 * n) return W_BAD;
  *count = (size_t)n;
  if (out) {
    if (n > cap) return W_SHORT;
    for (uint64_t k = 0; k < n; k++) {
      ofr x;
      memcpy(x.l, p + 8 + 32 * k, 32);
      if (ofr_geq_p(x.l)) return W_BAD; /* ark rejects non-canonical integers */
      out[k] = ofr_from_int(x);
    }
  }
  r->pos += nbytes;
  return W_OK;
}

/* SparseMatrix (sparse_matrix.rs:149-155 field order, 118-147 decode).
 * rs/cols/vals may be NULL (sizes only). */
static int rd_matrix(rd* r, const ofr* interner, size_t interner_len, size_t cap_rows,
                     size_t cap_nnz, uint32_t* rs, uint32_t* cols, ofr* vals, size_t* num_rows,
                     size_t* num_cols, size_t* nnz) {
  uint64_t nr, nc, n1, n2 = 0, n3 = 0;
  if (rd_usize(r, &nr) || rd_usize(r, &nc) || rd_usize(r, &n1)) return W_BAD;
  if (n1 > r->len - r->pos) return W_BAD; /* each element takes a byte at least */
  uint32_t* rows_tmp = (uint32_t*)malloc(sizeof(uint32_t) * (n1 ? n1 : 1));
  if (!rows_tmp) return W_BAD;
  int err = W_OK;
  for (uint64_t i = 0; i < n1 && !err; i++) err = rd_u32(r, &rows_tmp[i]);
  if (!err) err = rd_usize(r, &n2);
  if (!err && n2 > r->len - r->pos) err = W_BAD;
  uint32_t* deltas = NULL;
  if (!err) {
    deltas = (uint32_t*)malloc(sizeof(uint32_t) * (n2 ? n2 : 1));
    if (!deltas) err = W_BAD;
  }
  for (uint64_t i = 0; i < n2 && !err; i++) err = rd_u32(r, &deltas[i]);
  if (!err) err = rd_usize(r, &n3);
  if (!err && n3 > r->len - r->pos) err = W_BAD;
  /* structure: num_rows is the row-start count (SparseMatrix::new and grow
   * keep them equal; row_range panics past the vector) */
  if (!err && (nr != n1 || n2 != n3 || (n1 && rows_tmp[0] != 0) || (!n1 && n3))) err = W_BAD;
  for (uint64_t i = 1; i < n1 && !err; i++)
    if (rows_tmp[i] < rows_tmp[i - 1]) err = W_BAD;
  if (!err && n1 && rows_tmp[n1 - 1] > n3) err = W_BAD;
  if (!err) {
    *num_rows = (size_t)n1;
    *num_cols = (size_t)nc;
    *nnz = (size_t)n3;
  }
  uint64_t* ids = NULL;
  if (!err) {
    ids = (uint64_t*)malloc(sizeof(uint64_t) * (n3 ? n3 : 1));
    if (!ids) err = W_BAD;
  }
  for (uint64_t i = 0; i < n3 && !err; i++) {
    err = rd_usize(r, &ids[i]);
    if (!err && ids[i] >= interner_len) err = W_BAD;
  }
  /* decode_col_deltas (sparse_matrix.rs:118-147): per non-empty row, the
   * first column absolute, the rest running sums (u32, wrapping) */
  if (!err && rs) {
    if (n1 > cap_rows || n3 > cap_nnz) {
      err = W_SHORT;
    } else {
      size_t di = 0;
      for (uint64_t row = 0; row < n1 && !err; row++) {
        const uint64_t s = rows_tmp[row], e = row + 1 < n1 ? rows_tmp[row + 1] : n3;
        uint32_t prev = 0;
        for (uint64_t k = s; k < e; k++, di++) {
          const uint32_t c = k == s ? deltas[di] : prev + deltas[di];
          if (c >= nc) err = W_BAD;
          cols[k] = c;
          prev = c;
        }
      }
      for (uint64_t i = 0; i < n1; i++) rs[i] = rows_tmp[i];
      for (uint64_t i = 0; i < n3 && !err; i++) vals[i] = interner[ids[i]];
    }
  } else if (!err) {
    /* sizes only: still check the columns */
    size_t di = 0;
    for (uint64_t row = 0; row < n1 && !err; row++) {
      const uint64_t s = rows_tmp[row], e = row + 1 < n1 ? rows_tmp[row + 1] : n3;
      uint32_t prev = 0;
      for (uint64_t k = s; k < e; k++, di++) {
        const uint32_t c = k == s ? deltas[di] : prev + deltas[di];
        if (c >= nc) err = W_BAD;
        prev = c;
      }
    }
  }
  free(rows_tmp);
  free(deltas);
  free(ids);
  return err;
}

int or_varint_decode(const uint8_t* b, size_t len, int max_bytes, int bits, uint64_t* out,
                     size_t* used) {
  rd r = {b, len, 0};
  int e = rd_varint(&r, max_bytes, bits, out);
  *used = r.pos;
  return e;
}

int or_interner_decode(const uint8_t* blob, size_t len, ofr* out, size_t cap, size_t* count) {
  rd r = {blob, len, 0};
  return rd_interner(&r, out, cap, count);
}

int or_sparse_matrix_decode(const uint8_t* blob, size_t len, const ofr* interner,
                            size_t interner_len, size_t cap_rows, size_t cap_nnz, uint32_t* rs,
                            uint32_t* cols, ofr* vals, size_t* num_rows, size_t* num_cols,
                            size_t* nnz) {
  rd r = {blob, len, 0};
  return rd_matrix(&r, interner, interner_len, cap_rows, cap_nnz, rs, cols, vals, num_rows,
                   num_cols, nnz);
}

/* R1CS (r1cs.rs:59-66).  sizes[0] = num_public_inputs, sizes[1] = interner
 * length, then (num_rows, num_cols, nnz) of a, b, c.  With rs[i] == NULL only
 * the sizes are produced. */
int or_r1cs_decode(const uint8_t* blob, size_t len, size_t sizes[11], uint32_t* const rs[3],
                   uint32_t* const cols[3], ofr* const vals[3], const size_t caps[6]) {
  rd r = {blob, len, 0};
  uint64_t npub;
  if (rd_usize(&r, &npub)) return W_BAD;
  sizes[0] = (size_t)npub;
  const size_t at = r.pos;
  size_t ni;
  int e = rd_interner(&r, NULL, 0, &ni);
  if (e) return e;
  sizes[1] = ni;
  ofr* inter = (ofr*)malloc(sizeof(ofr) * (ni ? ni : 1));
  if (!inter) return W_BAD;
  r.pos = at;
  e = rd_interner(&r, inter, ni, &ni);
  for (int m = 0; m < 3 && !e; m++)
    e = rd_matrix(&r, inter, ni, caps ? caps[2 * m] : 0, caps ? caps[2 * m + 1] : 0, rs[m],
                  cols[m], vals[m], &sizes[2 + 3 * m], &sizes[3 + 3 * m], &sizes[4 + 3 * m]);
  free(inter);
  return e;
}
