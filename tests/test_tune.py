"""Tile table of the autotuner (host side, no device work): class lookup,
the tile height an instance gets, and the text file round trip."""
import numpy as np
import pytest

Z_REPS = (4, 8, 12, 16, 24, 32, 48, 64)
F_REPS = (1 << 12, 1 << 15, 1 << 17, 1 << 19, 1 << 21)


@pytest.fixture
def table(hip):
    """Snapshot of every class, restored afterwards (the table is global)."""
    snap = {(dt, z, f): hip.tune_get(dt, z, f)
            for dt in (np.float32, np.float64) for z in Z_REPS for f in F_REPS}
    yield hip
    for (dt, z, f), rows in snap.items():
        hip.tune_set(dt, z, f, rows)


def test_classes_and_rows(table):
    hip = table
    hip.tune_set(np.float32, 8, 1000 * 10, 0)
    assert hip.tile_rows(np.float32, 1000, 10, 8) == 64       # built-in rule
    other_z = hip.tune_get(np.float32, 9, 2000)
    other_f = hip.tune_get(np.float32, 8, 1 << 14)
    hip.tune_set(np.float32, 8, 1000 * 10, 16)
    assert hip.tune_get(np.float32, 5, 2000) == 16           # same class: z 5..8, x*y < 2^14
    assert hip.tune_get(np.float32, 9, 2000) == other_z      # next z class untouched
    assert hip.tune_get(np.float32, 8, 1 << 14) == other_f   # next size class untouched
    assert hip.tile_rows(np.float32, 1000, 10, 8) == 16
    assert hip.tile_rows(np.float64, 1000, 10, 8) == 64       # per dtype
    hip.tune_set(np.float32, 8, 1000 * 10, 128)
    assert hip.tile_rows(np.float32, 1000, 10, 8) == 128
    # rows are a multiple of 4 and cover x evenly
    hip.tune_set(np.float32, 8, 1000 * 10, 48)
    r = hip.tile_rows(np.float32, 1000, 10, 8)
    assert r % 4 == 0 and r >= 48 and -(-1000 // r) == -(-1000 // 48)


def test_invalid_entries(table):
    hip = table
    with pytest.raises(ValueError):
        hip.tune_set(np.float32, 8, 100, 3)
    with pytest.raises(ValueError):
        hip.tune_set(np.float32, 65, 100, 64)
    # deeper fibers run the wide pass on fixed 16 x 16 tiles
    assert hip.tile_rows(np.float32, 10, 10, 65) == 16
    assert hip.tile_rows(np.float32, 10, 10, 0) == 0


def test_save_load_roundtrip(table, tmp_path):
    hip = table
    hip.tune_set(np.float32, 30, 1 << 17, 32)
    hip.tune_set(np.float64, 60, 1 << 21, 256)
    path = tmp_path / "t.txt"
    hip.tune_save(str(path))
    text = path.read_text()
    assert "f32 5 2 32" in text and "f64 7 4 256" in text
    hip.tune_set(np.float32, 30, 1 << 17, 0)
    hip.tune_load(str(path))
    assert hip.tune_get(np.float32, 30, 1 << 17) == 32
    assert hip.tune_source() == str(path)
    bad = tmp_path / "bad.txt"
    bad.write_text("f32 9 0 64\n")
    with pytest.raises(ValueError):
        hip.tune_load(str(bad))
    assert hip.tune_get(np.float32, 30, 1 << 17) == 32   # a bad file changes nothing


def test_empty_instances_plan(table):
    # an empty shape (x or y = 0) has no tiles: the query answers 0 rather
    # than planning (the batch planner skips such instances itself)
    hip = table
    for dt in (np.float32, np.float64):
        assert hip.tile_rows(dt, 0, 5, 3) == 0
        assert hip.tile_rows(dt, 7, 0, 3) == 0
        assert hip.tile_rows(dt, 0, 0, 1) == 0
        assert hip.tile_rows(dt, 7, 5, 3) >= 1


def test_shipped_table(table):
    """The library loads tune/gfx950.txt by default (checked in a fresh
    process: the table is global); the large single-instance classes carry
    the heights measured for them (2048x2048x64 and x32: 128 rows, FP64
    1024x1024x32: 28 rows) and the headline classes keep 32."""
    import os
    import subprocess
    import sys
    hip = table
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(root, "bnmf-algs_amd", "tune", "gfx950.txt")
    env = {k: v for k, v in os.environ.items() if k != "BNMF_TUNE_FILE"}
    src = subprocess.run([sys.executable, "-c", "import bnmf_hip; print(bnmf_hip.tune_source())"],
                         cwd=os.path.join(root, "bnmf-algs_amd"), env=env, capture_output=True,
                         text=True, timeout=120).stdout.strip()
    assert os.path.realpath(src) == os.path.realpath(path), src
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    assert rows and all(r[0] in ("f32", "f64") and int(r[3]) % 4 == 0 and 4 <= int(r[3]) <= 256
                        for r in rows)
    hip.tune_load(path)
    assert hip.tile_rows(np.float32, 2048, 2048, 64) == 128
    assert hip.tile_rows(np.float32, 2048, 2048, 32) == 128
    assert hip.tile_rows(np.float64, 1024, 1024, 32) == 28
    assert hip.tune_get(np.float32, 27, 300 * 300) == 32


def test_large_lone_instances_keep_enough_tiles(table):
    """A cap shortens tiles but never lengthens them past what keeps the GPU
    filled: an instance of 2^20 fibers or more gets at least enough row tiles
    for 512 tiles in all (FP32, 64-fiber columns) or 1024 (FP64, 32-fiber
    columns at z <= 32), with tiles of at least 16 rows; smaller instances
    get ceil(x / cap) tiles of equal height, a multiple of 4."""
    hip = table
    # FP32 4096 x 256 x 8: 4 column tiles, so 128 row tiles of 32 rows even
    # under a 512-row cap
    hip.tune_set(np.float32, 8, 4096 * 256, 512)
    r = hip.tile_rows(np.float32, 4096, 256, 8)
    assert r == 32 and -(-4096 // r) * (256 // 64) >= 512
    # FP64 4096 x 256 x 8: 8 column tiles -> 128 row tiles of 32 rows
    hip.tune_set(np.float64, 8, 4096 * 256, 512)
    r = hip.tile_rows(np.float64, 4096, 256, 8)
    assert r == 32 and -(-4096 // r) * (256 // 32) >= 1024
    # a 1024-row cap on 1024 x 1024 FP32 (16 column tiles): 32 row tiles of
    # 32 rows
    hip.tune_set(np.float32, 8, 1 << 20, 1024)
    assert hip.tile_rows(np.float32, 1024, 1024, 8) == 32
    # below 2^20 fibers the cap alone decides: 1000 rows under 512 -> 2 x 500
    hip.tune_set(np.float32, 8, 1000 * 10, 512)
    assert hip.tile_rows(np.float32, 1000, 10, 8) == 500
