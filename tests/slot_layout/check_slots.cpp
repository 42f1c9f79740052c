// Checks the row-slot layout builder on the host: every row's slots, decoded
// with the same slot_decode the kernels use, must give the columns of its
// owner faces (face order) and neighbour faces (losort order), and, for the
// symmetric form, each neighbour slot's owner slot must hold the same face
// (except in tiles flagged kSlotNbrPlanes, which store those coefficients
// themselves).  Rows of face tiles have no slots and run the face loops;
// a tile becomes one only for a reason (a row of more than kShortRow faces
// a side, or offsets that no base reaches within the int16 deltas).
//
// Input (binary, int32): nMeshes, then per mesh n, nFaces, nTiles,
// tileStart[nTiles+1] when nTiles > 0 (else 256-row tiles from row 0),
// ownStart[n+1], losortStart[n+1], upperAddr[nFaces], lowerAddr[nFaces],
// losort[nFaces].
// Output: one line per mesh, "slots=<0|1> patterned=<rows> explicit=<rows>
// faces=<rows> nbrplanes=<rows> errors=<count> major=<offset> minor=<offset>"
// (the dominant strides).

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hipfoam_slots.hpp"

using namespace hipfoam;

static bool read_ints(FILE* f, std::vector<int32_t>& v, size_t count) {
  v.resize(count);
  return count == 0 || fread(v.data(), sizeof(int32_t), count, f) == count;
}

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: check_slots <addressing.bin>\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  int32_t nm = 0;
  if (fread(&nm, sizeof(int32_t), 1, f) != 1) return 2;
  constexpr int K = kShortRow, S = 2 * kShortRow;
  for (int m = 0; m < nm; ++m) {
    int32_t hdr[3];
    if (fread(hdr, sizeof(int32_t), 3, f) != 3) return 2;
    const int n = hdr[0], nf = hdr[1], nt = hdr[2];
    std::vector<int32_t> ts, own, los, upp, low, lso;
    if ((nt > 0 && !read_ints(f, ts, nt + 1)) || !read_ints(f, own, n + 1) ||
        !read_ints(f, los, n + 1) || !read_ints(f, upp, nf) || !read_ints(f, low, nf) ||
        !read_ints(f, lso, nf))
      return 2;
    // slot tile of every row
    std::vector<int> tileOf(n);
    for (int i = 0; i < n; ++i) tileOf[i] = i / kSlotTile;
    for (int t = 0; t < nt; ++t)
      for (int i = ts[t]; i < ts[t + 1]; ++i) tileOf[i] = t;
    SlotLayout L;
    const bool ok = build_slots(n, own.data(), upp.data(), los.data(), lso.data(), low.data(), L,
                                nt > 0 ? ts.data() : nullptr, nt);
    if (!ok) {
      printf("slots=0 patterned=0 explicit=0 faces=%d nbrplanes=0 errors=0 major=0 minor=0\n",
             n);
      continue;
    }
    long errors = 0;
    int64_t patterned = 0;
    auto decode = [&](int i, int (&o)[S], unsigned& ow) {
      const int32_t* d = &L.tile[(size_t)tileOf[i] * kSlotDesc];
      slot_decode(d, L.row[i], L.words.data(), (size_t)n, i, o, ow);
    };
    int64_t faceRows = 0, nbrRows = 0;
    // a face tile must have a reason: a long row, or offsets beyond any base
    // reachable with int16 deltas (more than 3 clusters of offsets 65535 wide)
    std::vector<int> tileLong(L.tile.size() / kSlotDesc, 0);
    for (int i = 0; i < n; ++i)
      if (own[i + 1] - own[i] > K || los[i + 1] - los[i] > K) tileLong[tileOf[i]] = 1;
    for (size_t t = 0; t < tileLong.size(); ++t) {
      if (!(L.tile[t * kSlotDesc + 6] & kSlotFaces) || tileLong[t]) continue;
      bool far = false;
      for (int i = 0; i < n && !far; ++i) {
        if (tileOf[i] != (int)t) continue;
        for (int k = own[i]; k < own[i + 1]; ++k) far = far || std::abs(upp[k] - i) > 32767;
        for (int k = los[i]; k < los[i + 1]; ++k) far = far || std::abs(low[lso[k]] - i) > 32767;
      }
      if (!far) ++errors;
    }
    for (int i = 0; i < n; ++i) {
      int o[S];
      unsigned ow;
      decode(i, o, ow);
      const int flags = L.tile[(size_t)tileOf[i] * kSlotDesc + 6];
      if (flags & kSlotFaces) {
        ++faceRows;
        for (int q = 0; q < S; ++q) errors += o[q] != 0;
        continue;
      }
      if (flags & kSlotNbrPlanes) ++nbrRows;
      if (flags & kSlotPatterned) ++patterned;
      // owner slots: the row's owner faces in face order
      int k = own[i];
      for (int q = 0; q < K; ++q)
        if (o[q] != 0 && (k >= own[i + 1] || i + o[q] != upp[k++])) ++errors;
      if (k != own[i + 1]) ++errors;
      // neighbour slots: the losort faces in order, each found again in the
      // owner slot its packed index names
      k = los[i];
      for (int q = K; q < S; ++q) {
        if (o[q] == 0) continue;
        if (k >= los[i + 1]) {
          ++errors;
          continue;
        }
        const int face = lso[k++];
        const int col = i + o[q];
        if (col != low[face] || col < 0 || col >= n) {
          ++errors;
          continue;
        }
        // the coefficient is this row's own (plane q) in kSlotNbrPlanes
        // tiles; elsewhere it is found in the owner slot the packed index names
        if (flags & kSlotNbrPlanes) continue;
        const int cflags = L.tile[(size_t)tileOf[col] * kSlotDesc + 6];
        if (cflags & kSlotFaces) {
          ++errors;  // an owner without slots needs kSlotNbrPlanes here
          continue;
        }
        const unsigned s = (ow >> (2 * (q - K))) & 3u;
        int oc[S];
        unsigned owc;
        decode(col, oc, owc);
        if (s >= (unsigned)K || oc[s] == 0 || col + oc[s] != i) {
          ++errors;
          continue;
        }
        // the face of owner slot s of row col: its rank among col's filled owner slots
        int rank = 0;
        for (int p = 0; p < (int)s; ++p) rank += oc[p] != 0;
        if (own[col] + rank != face) ++errors;
      }
      if (k != los[i + 1]) ++errors;
    }
    printf("slots=1 patterned=%lld explicit=%lld faces=%lld nbrplanes=%lld errors=%ld "
           "major=%lld minor=%lld\n",
           (long long)L.patternedRows, (long long)(n - L.patternedRows - faceRows),
           (long long)faceRows, (long long)nbrRows,
           errors + (patterned != L.patternedRows) + (faceRows != L.faceRows) +
               (nbrRows != L.nbrPlaneRows),
           (long long)L.strideMajor, (long long)L.strideMinor);
  }
  fclose(f);
  return 0;
}
