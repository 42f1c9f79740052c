// Row-slot layout of an LDU mesh (LduOp::slotTile, slotRow, slotOff): its
// constants, the per-row decode shared by the kernels and the host, and the
// host builder.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace hipfoam {

// Rows of hexahedral meshes have at most 3 owner and 3 neighbour faces.
constexpr int kShortRow = 3;
// Rows per offset base of the row-slot layout, and the empty-slot delta.
constexpr int kSlotTile = 256;
constexpr int kSlotEmpty = -32768;
constexpr int kSlotDesc = 8;       // int32 words per tile descriptor
constexpr int kSlotPatterned = 1;  // descriptor flag: every face at its slot base
// descriptor flag: a face tile, whose rows do not fit the slots (a row with
// more than kShortRow faces on a side, or offsets beyond the int16 deltas);
// its rows run the LDU face loops
constexpr int kSlotFaces = 2;
// descriptor flag: a neighbour face of a row of this tile is owned by a row of
// a face tile, which has no owner slots; symmetric operators then store this
// tile's neighbour coefficients in planes 3..5 of its own rows, as
// asymmetric ones do
constexpr int kSlotNbrPlanes = 4;

// Column offsets of row i's slots (0: an empty slot) from its tile's
// descriptor d and its slot byte rb, and the packed owner slots of the row's
// neighbour faces (used by symmetric operators).  Host-callable, so the layout
// builder's output can be checked against the face loops on the CPU.
__host__ __device__ __forceinline__ void slot_decode(const int32_t* d, unsigned rb,
                                                     const uint32_t* __restrict__ slotOff,
                                                     size_t n, int i,
                                                     int (&o)[2 * kShortRow], unsigned& own) {
  const int flags = d[6];
  if (flags & kSlotFaces) {
#pragma unroll
    for (int q = 0; q < 2 * kShortRow; ++q) o[q] = 0;
    own = 0;
  } else if (flags & kSlotPatterned) {
#pragma unroll
    for (int q = 0; q < 2 * kShortRow; ++q) o[q] = ((rb >> q) & 1u) ? d[q] : 0;
    own = (unsigned)flags >> 8;
  } else {
    uint32_t w[kShortRow];
#pragma unroll
    for (int p = 0; p < kShortRow; ++p) w[p] = slotOff[p * n + i];
#pragma unroll
    for (int p = 0; p < kShortRow; ++p) {
      const int d0 = (int)(int16_t)(w[p] & 0xffffu), d1 = (int)(int16_t)(w[p] >> 16);
      o[2 * p] = d0 == kSlotEmpty ? 0 : d[2 * p] + d0;
      o[2 * p + 1] = d1 == kSlotEmpty ? 0 : d[2 * p + 1] + d1;
    }
    own = rb;
  }
}

// Row-slot layout of a mesh (LduOp::slotTile, slotRow, slotOff), built on the
// host once per topology.
struct SlotLayout {
  std::vector<int32_t> tile;    // kSlotDesc words per tile
  std::vector<uint32_t> words;  // kShortRow * n: int16 delta pairs
  std::vector<uint8_t> row;     // n: presence mask or packed owner slots
  int64_t patternedRows = 0;    // rows in patterned tiles
  int64_t faceRows = 0;         // rows in face tiles (LDU face loops)
  int64_t nbrPlaneRows = 0;     // rows in tiles flagged kSlotNbrPlanes
  // the most common largest and second-largest owner offsets of the
  // patterned tiles (the z and y strides of a structured block), 0 if none
  int64_t strideMajor = 0, strideMinor = 0;
};

// Each row's owner faces go, in face order, into increasing slots among 0..2,
// and its neighbour faces, in losort order, into increasing slots among 3..5
// (a row with fewer faces leaves slots empty, in any position, so the face
// order and hence the summation order is kept).  Every kSlotTile-row tile has
// one column offset base per slot; a face is stored as the int16 delta of its
// offset (column - row) from its slot's base.
//
// Candidate bases of a tile side, tried in turn until every row of the tile
// fits: the side's distinct offsets when there are at most three (on a
// structured mesh these make every delta zero); the median of each position
// over the rows with the most faces on that side; zero; then the centres of
// the clusters of the tile's offsets (values within an int16 window), spread
// over the three slots in increasing order.  A row takes the
// placement of its faces that keeps its largest |delta| smallest.  A tile
// whose deltas are all zero, and whose neighbour faces of each slot all sit
// in the same owner slot of their owner rows, is patterned.  A tile with a
// row of more than kShortRow faces on a side, or where no candidate fits,
// becomes a face tile (its rows run the face loops); a slot tile with a
// neighbour face owned by a face-tile row is flagged kSlotNbrPlanes.  False
// when every tile is a face tile, or when a losort face is not among its
// owner's faces (inconsistent addressing).
// Tiles are kSlotTile-row runs from row 0, or, with tileStart, the ntiles
// runs [tileStart[t], tileStart[t + 1]) of at most kSlotTile rows (a batch's
// work tiles, which never straddle two systems).
inline bool build_slots(int n, const int32_t* ownStart, const int32_t* uppAddr,
                 const int32_t* losortStart, const int32_t* losort, const int32_t* lowAddr,
                 SlotLayout& L, const int32_t* tileStart = nullptr, int ntiles = 0) {
  constexpr int K = kShortRow, S = 2 * kShortRow;
  constexpr int64_t kHalf = 32767;
  const int nt = tileStart ? ntiles : (n + kSlotTile - 1) / kSlotTile;
  // rows [i0, i1) of tile t
  auto tile_rows = [&](int t, int& i0, int& i1) {
    if (tileStart) {
      i0 = tileStart[t];
      i1 = tileStart[t + 1];
    } else {
      i0 = t * kSlotTile;
      i1 = std::min(n, i0 + kSlotTile);
    }
  };
  L.tile.assign((size_t)nt * kSlotDesc, 0);
  L.words.assign((size_t)K * n, 0);
  L.row.assign((size_t)n, 0);
  L.patternedRows = L.faceRows = L.nbrPlaneRows = 0;
  L.strideMajor = L.strideMinor = 0;
  std::vector<int16_t> delta((size_t)S * n, (int16_t)kSlotEmpty);
  // offsets of row i on one side (owner: side 0, neighbour: side 1), in face order
  auto offsets = [&](int i, int side, int64_t* off) -> int {
    if (side == 0) {
      const int f0 = ownStart[i], cnt = ownStart[i + 1] - f0;
      for (int q = 0; q < cnt && q < K; ++q) off[q] = (int64_t)uppAddr[f0 + q] - i;
      return cnt;
    }
    const int k0 = losortStart[i], cnt = losortStart[i + 1] - k0;
    for (int q = 0; q < cnt && q < K; ++q) off[q] = (int64_t)lowAddr[losort[k0 + q]] - i;
    return cnt;
  };
  // best increasing placement of cnt offsets on bases bs; false if it does not fit
  auto place = [&](const int64_t* off, int cnt, const int64_t* bs, int* pos) -> bool {
    int64_t bestErr = INT64_MAX;
    for (int mask = 0; mask < (1 << K); ++mask) {
      if (__builtin_popcount(mask) != cnt) continue;
      int ps[K], m = 0;
      for (int q = 0; q < K; ++q)
        if (mask >> q & 1) ps[m++] = q;
      int64_t err = 0;
      for (int j = 0; j < cnt; ++j) {
        const int64_t e = off[j] - bs[ps[j]];
        err = std::max(err, e < 0 ? -e : e);
      }
      if (err < bestErr) {
        bestErr = err;
        for (int j = 0; j < cnt; ++j) pos[j] = ps[j];
      }
    }
    return cnt == 0 || bestErr <= kHalf;
  };
  std::vector<int64_t> most[K], all;
  for (int q = 0; q < K; ++q) most[q].reserve(kSlotTile);
  all.reserve((size_t)K * kSlotTile);
  // places every row of the tile side on bases bs; false if one does not fit.
  // zero: every delta of the side is zero.
  auto try_bases = [&](int i0, int i1, int side, const int64_t* bs, bool& zero) -> bool {
    zero = true;
    for (int i = i0; i < i1; ++i) {
      int64_t off[K];
      int pos[K];
      const int cnt = offsets(i, side, off);
      if (!place(off, cnt, bs, pos)) return false;
      int16_t* d = &delta[(size_t)i * S + side * K];
      for (int q = 0; q < K; ++q) d[q] = (int16_t)kSlotEmpty;
      for (int j = 0; j < cnt; ++j) {
        d[pos[j]] = (int16_t)(off[j] - bs[pos[j]]);
        zero = zero && d[pos[j]] == 0;
      }
    }
    return true;
  };
  std::vector<uint8_t> zeroTile((size_t)nt, 1), faceTile((size_t)nt, 0);
  for (int t = 0; t < nt; ++t) {
    int i0, i1;
    tile_rows(t, i0, i1);
    bool faces = false;
    for (int side = 0; side < 2 && !faces; ++side) {
      for (int q = 0; q < K; ++q) most[q].clear();
      all.clear();
      int mcnt = 0;
      for (int i = i0; i < i1 && !faces; ++i) {
        int64_t off[K];
        const int cnt = offsets(i, side, off);
        if (cnt > K) {
          faces = true;
          break;
        }
        if (cnt > mcnt) {
          mcnt = cnt;
          for (int q = 0; q < K; ++q) most[q].clear();
        }
        for (int q = 0; q < cnt; ++q) {
          all.push_back(off[q]);
          if (cnt == mcnt) most[q].push_back(off[q]);
        }
      }
      // candidates in order; the first that fits with zero deltas wins, else
      // the first that fits
      std::vector<std::array<int64_t, K>> cand;
      std::sort(all.begin(), all.end());
      const size_t ndist = std::unique(all.begin(), all.end()) - all.begin();
      if (ndist > 0 && ndist <= (size_t)K) {
        std::array<int64_t, K> dist{0, 0, 0};
        std::copy(all.begin(), all.begin() + ndist, dist.begin());
        cand.push_back(dist);
      }
      if (mcnt > 0) {
        std::array<int64_t, K> med{0, 0, 0};
        for (int q = 0; q < mcnt; ++q) {
          std::nth_element(most[q].begin(), most[q].begin() + most[q].size() / 2,
                           most[q].end());
          med[q] = most[q][most[q].size() / 2];
        }
        cand.push_back(med);
      }
      cand.push_back({0, 0, 0});
      all.resize(ndist);
      int64_t centre[K + 1];
      int nc = 0;
      for (size_t a = 0; a < all.size() && nc <= K;) {
        size_t e = a;
        while (e + 1 < all.size() && all[e + 1] - all[a] <= 2 * kHalf) ++e;
        centre[nc++] = all[a] + (all[e] - all[a]) / 2;
        a = e + 1;
      }
      if (nc == 1) {
        cand.push_back({centre[0], centre[0], centre[0]});
      } else if (nc == 2) {
        cand.push_back({centre[0], centre[1], centre[1]});
        cand.push_back({centre[0], centre[0], centre[1]});
      } else if (nc == 3) {
        cand.push_back({centre[0], centre[1], centre[2]});
      }
      if (faces) break;
      int chosen = -1;
      bool chosenZero = false;
      for (size_t c = 0; c < cand.size(); ++c) {
        bool zero = false;
        if (!try_bases(i0, i1, side, cand[c].data(), zero)) continue;
        if (chosen < 0 || zero) {
          chosen = (int)c;
          chosenZero = zero;
        }
        if (zero) break;
      }
      if (chosen < 0) {
        faces = true;
        break;
      }
      for (int q = 0; q < K; ++q) {
        const int64_t b = cand[chosen][q];
        if (b < INT32_MIN || b > INT32_MAX) faces = true;
      }
      if (faces) break;
      // the rows hold the last placement tried: place them on the chosen bases again
      bool zero = false;
      try_bases(i0, i1, side, cand[chosen].data(), zero);
      for (int q = 0; q < K; ++q)
        L.tile[(size_t)t * kSlotDesc + side * K + q] = (int32_t)cand[chosen][q];
      zeroTile[t] = zeroTile[t] && chosenZero;
    }
    if (faces) {
      // no slots in this tile: its rows run the face loops
      faceTile[t] = 1;
      zeroTile[t] = 0;
      for (int q = 0; q < kSlotDesc; ++q) L.tile[(size_t)t * kSlotDesc + q] = 0;
      for (int i = i0; i < i1; ++i)
        for (int q = 0; q < S; ++q) delta[(size_t)i * S + q] = (int16_t)kSlotEmpty;
    }
  }
  int64_t nFaceTiles = 0;
  for (int t = 0; t < nt; ++t) nFaceTiles += faceTile[t];
  if (nt > 0 && nFaceTiles == nt) return false;
  // slot tile of every row
  std::vector<int32_t> tileOf((size_t)n, 0);
  for (int t = 0; t < nt; ++t) {
    int i0, i1;
    tile_rows(t, i0, i1);
    for (int i = i0; i < i1; ++i) tileOf[i] = t;
  }
  // owner slot of every face, then each row's packed owner slots of its
  // neighbour faces (2 bits per neighbour slot)
  std::vector<int8_t> ownSlot((size_t)(n > 0 ? ownStart[n] : 0), -1);
  for (int i = 0; i < n; ++i) {
    int f = ownStart[i];
    for (int q = 0; q < K; ++q)
      if (delta[(size_t)i * S + q] != kSlotEmpty) ownSlot[f++] = (int8_t)q;
  }
  std::vector<uint8_t> packed((size_t)n, 0), nbrPlanes((size_t)nt, 0);
  std::map<std::pair<int64_t, int64_t>, int64_t> strides;
  for (int i = 0; i < n; ++i) {
    int k = losortStart[i];
    unsigned pk = 0;
    for (int q = K; q < S; ++q) {
      if (delta[(size_t)i * S + q] == kSlotEmpty) continue;
      const int f = losort[k++];
      const int col = lowAddr[f];
      if (col < 0 || col >= n || f < ownStart[col] || f >= ownStart[col + 1]) return false;
      if (ownSlot[f] < 0) {
        // owned by a face-tile row: the coefficient is stored in this row
        nbrPlanes[tileOf[i]] = 1;
        continue;
      }
      pk |= (unsigned)ownSlot[f] << (2 * (q - K));
    }
    packed[i] = (uint8_t)pk;
  }
  for (int t = 0; t < nt; ++t) {
    int i0, i1;
    tile_rows(t, i0, i1);
    if (faceTile[t]) {
      L.tile[(size_t)t * kSlotDesc + 6] = kSlotFaces;
      L.faceRows += i1 - i0;
      continue;
    }
    if (nbrPlanes[t]) L.nbrPlaneRows += i1 - i0;
    // patterned: zero deltas, and one owner slot per neighbour slot (the
    // latter only where symmetric operators gather from the owner rows)
    bool pat = zeroTile[t];
    int ownQ[K] = {-1, -1, -1};
    for (int i = i0; i < i1 && pat && !nbrPlanes[t]; ++i)
      for (int q = K; q < S && pat; ++q) {
        if (delta[(size_t)i * S + q] == kSlotEmpty) continue;
        const int s = (packed[i] >> (2 * (q - K))) & 3;
        if (ownQ[q - K] < 0) ownQ[q - K] = s;
        pat = ownQ[q - K] == s;
      }
    if (!pat) L.tile[(size_t)t * kSlotDesc + 6] = nbrPlanes[t] ? kSlotNbrPlanes : 0;
    if (pat) {
      unsigned pk = 0;
      for (int q = 0; q < K; ++q) pk |= (unsigned)std::max(ownQ[q], 0) << (2 * q);
      L.tile[(size_t)t * kSlotDesc + 6] =
          kSlotPatterned | (nbrPlanes[t] ? kSlotNbrPlanes : 0) | (int32_t)(pk << 8);
      L.patternedRows += i1 - i0;
      // the tile's two largest owner offsets
      int64_t o[K];
      for (int q = 0; q < K; ++q) o[q] = L.tile[(size_t)t * kSlotDesc + q];
      std::sort(o, o + K);
      const std::pair<int64_t, int64_t> key(std::max<int64_t>(o[K - 1], 0),
                                            std::max<int64_t>(o[K - 2], 0));
      ++strides[key];
    }
    for (int i = i0; i < i1; ++i) {
      const int16_t* d = &delta[(size_t)i * S];
      unsigned mask = 0;
      for (int q = 0; q < S; ++q)
        if (d[q] != kSlotEmpty) mask |= 1u << q;
      L.row[i] = pat ? (uint8_t)mask : packed[i];
      for (int p = 0; p < K; ++p)
        L.words[(size_t)p * n + i] =
            (uint32_t)(uint16_t)d[2 * p] | ((uint32_t)(uint16_t)d[2 * p + 1] << 16);
    }
  }
  int64_t bestCount = 0;
  for (const auto& kv : strides)
    if (kv.second > bestCount) {
      bestCount = kv.second;
      L.strideMajor = kv.first.first;
      L.strideMinor = kv.first.second;
    }
  return true;
}

}  // namespace hipfoam
