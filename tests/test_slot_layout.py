"""Row-slot layout builder (host code of the HIP library), checked on the CPU.

tests/slot_layout/check_slots.cpp runs the library's own layout builder and
its per-row decode (the function the kernels call) over each mesh and
counts rows whose decoded slots differ from the face loops: owner faces in
face order, neighbour faces in losort order, and each neighbour slot's packed
owner slot pointing at the same face.  Structured blocks must come out fully
patterned (1 byte of addressing per row).  A tile holding a row of more than
three faces a side, or offsets no base reaches, becomes a face tile whose
rows run the face loops, with the rest of the mesh kept on slots; only a
mesh whose every tile is a face tile is refused (the library then runs the
face loops everywhere)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from metal4foam_amd import meshgen as mg

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "metal4foam_amd", "csrc")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    return None


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cc = hipcc()
    if cc is None:
        pytest.skip("hipcc not found")
    out = str(tmp_path_factory.mktemp("slots") / "check_slots")
    subprocess.run([cc, "--offload-host-only", "-O2", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "slot_layout", "check_slots.cpp"), "-o", out],
                   check=True)
    return out


def from_pairs(n, pairs):
    pairs = sorted({(min(a, b), max(a, b)) for a, b in pairs if a != b})
    low = np.array([p[0] for p in pairs], dtype=np.int32)
    upp = np.array([p[1] for p in pairs], dtype=np.int32)
    own, los, loss = mg.derived_addressing(n, low, upp)
    return n, low, upp, own, los, loss


def block(nx, ny, nz):
    return mg.structured_addressing(nx, ny, nz)


def concat(blocks):
    n0, lows, upps = 0, [], []
    for n, low, upp, *_ in blocks:
        lows.append(low + n0)
        upps.append(upp + n0)
        n0 += n
    low = np.concatenate(lows).astype(np.int32)
    upp = np.concatenate(upps).astype(np.int32)
    own, los, loss = mg.derived_addressing(n0, low, upp)
    return n0, low, upp, own, los, loss


def work_tiles(mesh, sizes):
    starts, r0 = [], 0
    for s in sizes:
        starts += list(range(r0, r0 + s, 256))
        r0 += s
    return (*mesh, starts + [r0])


def renumbered(nx, ny, nz, seed):
    n, low, upp, *_ = block(nx, ny, nz)
    perm = np.random.default_rng(seed).permutation(n)
    return from_pairs(n, zip(perm[low].tolist(), perm[upp].tolist()))


def irregular(n, seed):
    rng = np.random.default_rng(seed)
    pairs = [(i, i + 1) for i in range(n - 1)]
    pairs += [tuple(rng.choice(n, 2, replace=False)) for _ in range(n)]
    return from_pairs(n, pairs)


def scattered(n, seed):
    perm = np.random.default_rng(seed).permutation(n)
    pairs = [(i, i + 1) for i in range(n - 1)]
    pairs += [(int(perm[k]), int(perm[k + 1])) for k in range(0, n - 1, 2)]
    return from_pairs(n, pairs)


def block_with_long_rows(nx, ny, nz, extra, seed):
    """A structured block plus `extra` faces between random far-apart cells:
    their rows have more than three faces a side (face tiles) and their
    neighbours' tiles refer to owners without slots."""
    n, low, upp, *_ = block(nx, ny, nz)
    rng = np.random.default_rng(seed)
    pairs = list(zip(low.tolist(), upp.tolist()))
    for _ in range(extra):
        a = int(rng.integers(0, n))
        for d in (1, 2, 3):
            pairs.append((a, (a + d * 4099) % n))
    return from_pairs(n, pairs)


def face_order_shuffled(nx, ny, nz, seed):
    """A structured block whose faces of one owner are not in neighbour order."""
    n, low, upp, *_ = block(nx, ny, nz)
    rng = np.random.default_rng(seed)
    order = np.lexsort((rng.random(low.shape[0]), low))
    low, upp = low[order].astype(np.int32), upp[order].astype(np.int32)
    own, los, loss = mg.derived_addressing(n, low, upp)
    return n, low, upp, own, los, loss


CASES = {
    "box": (lambda: block(33, 21, 17), "all"),
    "plane": (lambda: block(100, 100, 1), "all"),
    "wide": (lambda: block(200, 170, 3), "all"),
    "thin": (lambda: block(7, 5, 300), "all"),
    "line": (lambda: block(1000, 1, 1), "all"),
    "blocks": (lambda: concat([block(20, 20, 20), block(23, 23, 23), block(17, 19, 21)]),
               "mixed"),
    # the same mesh on a batch's work tiles (256-row runs restarting at every
    # block): no tile straddles two blocks, so every tile is patterned
    "blocks_work_tiles": (lambda: work_tiles(
        concat([block(20, 20, 20), block(23, 23, 23), block(17, 19, 21)]),
        [8000, 12167, 6783]), "all"),
    "face_order": (lambda: face_order_shuffled(9, 8, 7, 3), "any"),
    "renumbered": (lambda: renumbered(12, 11, 10, 4), "any"),
    "scattered": (lambda: scattered(3000, 5), "any"),
    "irregular": (lambda: irregular(2000, 6), "refused"),
    # random cell numbering of a hex mesh: owner rows of up to six faces and
    # offsets across the whole mesh: every tile holds a long row, so the whole
    # mesh runs the face loops
    "renumbered_hex": (lambda: renumbered(40, 40, 40, 7), "refused"),
    # a structured block with a few long rows: a few face tiles, the rest on
    # slots (patterned where nothing refers to a face tile)
    "long_rows": (lambda: block_with_long_rows(44, 42, 40, 4, 8), "some_faces"),
    "one_cell": (lambda: block(1, 1, 1), "all"),
}


# dominant (largest, second-largest) owner offsets: the z and y strides
STRIDES = {"box": (33 * 21, 33), "plane": (100, 1), "wide": (200 * 170, 200),
           "thin": (35, 7), "line": (1, 0), "one_cell": (0, 0),
           "blocks_work_tiles": (23 * 23, 23)}


def run_checker(checker, tmp_path, meshes):
    """meshes: addressing tuples, each optionally followed by its tile starts."""
    path = tmp_path / "addr.bin"
    with open(path, "wb") as f:
        np.array([len(meshes)], dtype=np.int32).tofile(f)
        for mesh in meshes:
            n, low, upp, own, los, loss = mesh[:6]
            tiles = mesh[6] if len(mesh) > 6 else None
            nt = 0 if tiles is None else len(tiles) - 1
            np.array([n, low.shape[0], nt], dtype=np.int32).tofile(f)
            if tiles is not None:
                np.asarray(tiles, dtype=np.int32).tofile(f)
            for a in (own, loss, upp, low, los):
                np.ascontiguousarray(a, dtype=np.int32).tofile(f)
    out = subprocess.run([checker, str(path)], check=True, capture_output=True, text=True)
    rows = []
    for line in out.stdout.strip().splitlines():
        rows.append({k: int(v) for k, v in (kv.split("=") for kv in line.split())})
    return rows


def test_slot_layout_matches_face_loops(checker, tmp_path):
    names = list(CASES)
    meshes = [CASES[k][0]() for k in names]
    results = run_checker(checker, tmp_path, meshes)
    assert len(results) == len(names)
    for name, mesh, res in zip(names, meshes, results):
        expect = CASES[name][1]
        n = mesh[0]
        print(name, n, res)
        assert res["errors"] == 0, name
        if expect == "refused":
            assert res["slots"] == 0, name
            continue
        if expect != "any":
            assert res["slots"] == 1, name
        if expect == "all":
            assert res["patterned"] == n, name
        if name in STRIDES:
            assert (res["major"], res["minor"]) == STRIDES[name], name
        if expect == "mixed":
            assert 0 < res["explicit"] < n // 10, name
            assert res["patterned"] + res["explicit"] == n
        if expect in ("all", "mixed"):
            assert res["faces"] == 0, name
        if expect == "faces":
            assert res["faces"] > n // 2, name
        if expect == "some_faces":
            assert 0 < res["faces"] < n // 10, name
            assert res["nbrplanes"] > 0 and res["patterned"] > 0.8 * n, name
            assert res["patterned"] + res["explicit"] + res["faces"] == n


def test_slot_layout_empty_mesh(checker, tmp_path):
    z = np.zeros(0, dtype=np.int32)
    res = run_checker(checker, tmp_path, [(0, z, z, np.zeros(1, np.int32), z,
                                           np.zeros(1, np.int32))])
    assert res == [{"slots": 1, "patterned": 0, "explicit": 0, "faces": 0, "nbrplanes": 0,
                    "errors": 0, "major": 0, "minor": 0}]
