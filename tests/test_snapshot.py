"""NICAM snapshot reader/writer (nsw6_snapshot_read/write, nsw6_verify_read/write) and
the stand-alone run on a snapshot directory.

* CPU: the written file set is exactly the one src/main.cpp reads (names from
  tests/golden/snapshot_files.json, parsed from the reference); every file has the
  size and index order of its reader in src/data_io.cpp:207-402 (checked against
  numpy reshapes of the raw bytes, including the [2][l][1][ij] precipitation files);
  read(write(state)) is the identity; the reference's own vgrid files are read back
  bit for bit; missing, short and long files are errors.
* GPU: run() on a snapshot equals the direct step, and its check table against a
  verification set written from the oracle stays within the parity tolerance.
"""
import json
import os
import shutil

import numpy as np
import pytest

from nicam_kokkos_microphysics_amd import _native as N
from nicam_kokkos_microphysics_amd import fields as F
from nicam_kokkos_microphysics_amd import snapshot as S

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "snapshot_files.json")) as fh:
    FILES = json.load(fh)["files"]
# snapshot name -> state key
NAME_TO_KEY = {"rhogq_Lswp": "rhogq", "q_Lswp": "q", "precip_mp": "precip", "precip1_mp": "ISO1_precip",
               "precip2_mp": "ISO2_precip", "rhoein_precip_mp": "precip_rhoe", "lh_precip_mp": "precip_lh_heat",
               "rhophi_precip_mp": "precip_rhophi", "rhokin_precip_mp": "precip_rhokin", "qke": "rhogqke"}
LALL, IJ = 2, 40


def _state():
    gz, gzh, st = F.synth_state(LALL, IJ, seed=5, frac_towers=0.2)
    rng = np.random.default_rng(3)
    # every array distinct and nonzero, so a swapped file or axis cannot pass
    for k in st:
        st[k][...] = rng.standard_normal(st[k].shape)
    return gz, gzh, st


def _data_names(d):
    return sorted(os.path.splitext(n)[0] for n in os.listdir(os.path.join(d, "data")) if n.endswith(".dat"))


def test_file_set_is_the_reference_set(tmp_path):
    gz, gzh, st = _state()
    S.write(str(tmp_path), st, gz, gzh)
    S.write_verify(str(tmp_path), st)
    want_data = sorted(f["name"] for f in FILES if f["dir"] == "data")
    want_verify = sorted(f["name"] for f in FILES if f["dir"] == "ref_verify")
    assert _data_names(tmp_path) == want_data
    got_verify = sorted(os.path.splitext(n)[0] for n in os.listdir(tmp_path / "ref_verify"))
    assert got_verify == want_verify
    assert sorted(os.listdir(tmp_path / "data" / "vgrid")) == ["GRD_gz.dat", "GRD_gzh.dat"]


def test_file_shapes_follow_the_reference_readers(tmp_path):
    gz, gzh, st = _state()
    S.write(str(tmp_path), st, gz, gzh)
    for f in FILES:
        if f["dir"] != "data":
            continue
        raw = np.fromfile(tmp_path / "data" / (f["name"] + ".dat"), dtype="<f8")
        key = NAME_TO_KEY.get(f["name"], f["name"])
        if f["rank"] == 2:      # read_data_2d: [l][ij]
            np.testing.assert_array_equal(raw.reshape(LALL, IJ), st[key])
        elif f["rank"] == 4 and key in ("rhogq", "q"):   # [l][6][96][ij]
            np.testing.assert_array_equal(raw.reshape(LALL, 6, 96, IJ), st[key])
        elif f["rank"] == 4:    # View4D(2, lall, 1, ij): precipitation row slowest
            np.testing.assert_array_equal(raw.reshape(2, LALL, 1, IJ)[:, :, 0, :].transpose(1, 0, 2), st[key])
        elif st[key].ndim == 2:  # read_data_3d into (lall, 1, ij)
            np.testing.assert_array_equal(raw.reshape(LALL, 1, IJ)[:, 0, :], st[key])
        else:                    # read_data_3d: [l][96][ij]
            np.testing.assert_array_equal(raw.reshape(LALL, 96, IJ), st[key])
    np.testing.assert_array_equal(np.fromfile(tmp_path / "data" / "vgrid" / "GRD_gz.dat", "<f8"), gz)


def test_round_trip(tmp_path):
    gz, gzh, st = _state()
    S.write(str(tmp_path), st, gz, gzh)
    gz2, gzh2, back = S.read(str(tmp_path), LALL, IJ)
    np.testing.assert_array_equal(gz2, gz)
    np.testing.assert_array_equal(gzh2, gzh)
    in_files = {NAME_TO_KEY.get(f["name"], f["name"]) for f in FILES if f["dir"] == "data"}
    for k in st:
        if k in in_files:
            np.testing.assert_array_equal(back[k], st[k], err_msg=k)
        else:   # outputs the run does not read start at zero
            assert not back[k].any(), k
    S.write_verify(str(tmp_path), st)
    ver = S.read_verify(str(tmp_path), LALL, IJ)
    for k in S.VERIFY_KEYS:
        np.testing.assert_array_equal(ver[k], F.field_view(st, k))
    lines = S.verify_table(st, ver)
    assert len(lines) == 12 and lines[0].startswith("Checking [rhog] Max = 0.0000000000000000e+00")


def test_reference_vgrid_files_read_back(tmp_path):
    os.makedirs(tmp_path / "data" / "vgrid")
    for n in ("GRD_gz.dat", "GRD_gzh.dat"):
        shutil.copyfile(os.path.join(HERE, "golden", n), tmp_path / "data" / "vgrid" / n)
    _, _, st = _state()
    S.write(str(tmp_path), st)           # no grid: the vgrid files stay as copied
    gz, gzh, _ = S.read(str(tmp_path), LALL, IJ)
    np.testing.assert_array_equal(gz, np.fromfile(os.path.join(HERE, "golden", "GRD_gz.dat"), "<f8"))
    np.testing.assert_array_equal(gzh, np.fromfile(os.path.join(HERE, "golden", "GRD_gzh.dat"), "<f8"))


def test_missing_short_and_long_files_are_errors(tmp_path):
    gz, gzh, st = _state()
    S.write(str(tmp_path), st, gz, gzh)
    with pytest.raises(S.SnapshotError) as e:       # wrong column count
        S.read(str(tmp_path), LALL, IJ + 1)
    assert e.value.code == N.NSW6_IO_ERROR
    p = tmp_path / "data" / "tem.dat"
    raw = p.read_bytes()
    p.write_bytes(raw[:-8])
    with pytest.raises(S.SnapshotError):
        S.read(str(tmp_path), LALL, IJ)
    p.write_bytes(raw + raw[:8])
    with pytest.raises(S.SnapshotError):
        S.read(str(tmp_path), LALL, IJ)
    p.unlink()
    with pytest.raises(S.SnapshotError):
        S.read(str(tmp_path), LALL, IJ)
    with pytest.raises(S.SnapshotError):
        S.read_verify(str(tmp_path), LALL, IJ)      # no ref_verify/ written


@pytest.mark.gpu
def test_gpu_run_on_a_snapshot(tmp_path):
    import torch
    from oracle import oracle as O
    from nicam_kokkos_microphysics_amd.driver import NSW6
    gz, gzh, st = F.synth_state(1, 1200, seed=9, frac_supersat=0.05, frac_towers=0.2)
    S.write(str(tmp_path), st, gz, gzh)
    ref = F.copy_state(st)
    for _ in range(2):
        assert O.mp_driver(ref, gz, gzh, S.TIME_DTL).status == 0
    S.write_verify(str(tmp_path), ref)
    out, lines = S.run(str(tmp_path), lall=1, ijdim=1200, iterations=2)
    # the same as stepping the state directly
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, max_columns=1200)
    for _ in range(2):
        drv.step(dev, S.TIME_DTL, l_region=S.SET_L)
    torch.cuda.synchronize()
    direct = F.to_host(dev)
    drv.close()
    for k in F.OUTPUTS:
        np.testing.assert_array_equal(out[k], direct[k], err_msg=k)
    # check table against the oracle's verification set: within 1e-10 of max-abs
    assert len(lines) == 12
    for key, line in zip(S.VERIFY_KEYS, lines):
        mx = float(line.split("Max = ")[1].split(";")[0])
        mn = float(line.split("Min = ")[1].split(";")[0])
        scale = np.abs(F.field_view(ref, key)).max()
        assert max(abs(mx), abs(mn)) <= 1e-10 * scale, line
