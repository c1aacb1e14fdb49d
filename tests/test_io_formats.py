"""File formats of the reference's drivers (CPU): result/snapshot npz, gridded
background npy+json, observation cache with its SHA1 tag, background config
keys."""
import hashlib
import json
import os

import numpy as np
import pytest

from mun_amd import io as sem_io


def test_forward_results_npz_roundtrip(tmp_path):
    res = {"receiver_data": np.arange(12.0).reshape(4, 3), "dt": 2.5e-4, "nt": 4,
           "receiver_coords": [(0.0, 1.0), (2.0, 3.0), (4.0, 5.0)]}
    p = sem_io.save_forward_results(str(tmp_path), res)
    assert os.path.basename(p) == "sem_forward_results.npz"
    with np.load(p, allow_pickle=False) as f:
        assert set(f.files) == {"receiver_data", "dt", "nt", "receiver_coords"}
    back = sem_io.load_forward_results(p)
    assert back["dt"] == 2.5e-4 and back["nt"] == 4
    np.testing.assert_array_equal(back["receiver_coords"], np.asarray(res["receiver_coords"]))


def test_wavefield_snapshots_npz(tmp_path):
    snaps = [np.full(5, k, dtype=np.float64) for k in range(3)]
    p = sem_io.save_wavefield_snapshots(str(tmp_path), snaps, 1e-3, 600, -1.0, 1.0, -2.0, 2.0)
    with np.load(p, allow_pickle=False) as f:
        assert f["snapshots"].dtype == np.float32 and f["snapshots"].shape == (3, 5)
        assert float(f["zmax"]) == 2.0 and int(f["nt"]) == 600
    assert sem_io.save_wavefield_snapshots(str(tmp_path / "x"), [], 1e-3, 1, 0, 1, 0, 1) is None


def test_vp_grid_roundtrip(tmp_path):
    x = np.linspace(-500.0, 1500.0, 41)
    z = np.linspace(-100.0, 900.0, 21)
    vp = 1500.0 + np.add.outer(z, x) * 0.1
    npy, js = sem_io.save_vp_grid(str(tmp_path), "vp_background_pml", vp, x, z, "test",
                                  {"physical_subdomain": {"xmin": 0.0}})
    meta = json.load(open(js))
    assert meta["array_file"] == "vp_background_pml.npy"
    assert meta["grid"] == {"xmin": -500.0, "xmax": 1500.0, "zmin": -100.0, "zmax": 900.0, "dx": 50.0,
                            "dz": 50.0, "nx": 41, "nz": 21}
    for path in (js, npy):
        x2, z2, v2 = sem_io.load_vp_grid(path)
        np.testing.assert_allclose(x2, x)
        np.testing.assert_allclose(z2, z)
        np.testing.assert_array_equal(v2, vp)
    with pytest.raises(ValueError):
        sem_io.save_vp_grid(str(tmp_path), "bad", vp.T, x, z)


def test_background_config_keys_and_flips(tmp_path):
    vp = np.arange(12.0).reshape(3, 4)
    np.save(tmp_path / "vp.npy", vp)
    np.save(tmp_path / "x.npy", np.array([30.0, 20.0, 10.0, 0.0]))
    xg, zg, v = sem_io.background_from_config({"vp_path": str(tmp_path / "vp.npy"),
                                               "x_path": str(tmp_path / "x.npy"), "zmin": 5.0, "zmax": 25.0},
                                              -1.0, 1.0, -1.0, 1.0)
    assert xg.tolist() == [0.0, 10.0, 20.0, 30.0] and zg.tolist() == [5.0, 15.0, 25.0]
    assert v[0].tolist() == [3.0, 2.0, 1.0, 0.0]
    # metadata fallback: the domain extent with the array's sizes
    xg, zg, _ = sem_io.background_from_config({"vp": vp}, -1.0, 2.0, 0.0, 4.0)
    assert xg.tolist() == [-1.0, 0.0, 1.0, 2.0] and zg.tolist() == [0.0, 2.0, 4.0]
    with pytest.raises(ValueError):
        sem_io.background_from_config({}, 0, 1, 0, 1)


def test_observation_tag_and_cache(tmp_path):
    obj = {"b": [1, 2], "a": 0.5}
    payload = '{"a":0.5,"b":[1,2]}'
    assert sem_io.stable_json_hash(obj) == hashlib.sha1(payload.encode()).hexdigest()[:12]
    tag = sem_io.observation_tag("/data/vp_background_pml.npy", [(0, 10)] * 3, [(1, 2)] * 50, 1.2, 4e-4, 10.0,
                                 1.0, 0.01, 300.0, 4, (120, 60), 10.0, 1800.0, "smooth", np.zeros((6, 2)))
    assert tag.startswith("semtrue_vp_background_pml_3s_50r_T1.2_dt0.0004_f10_a1_noise0.01_")
    assert len(tag.rsplit("_", 1)[1]) == 12
    assert sem_io.load_observations(str(tmp_path), tag) is None
    noisy, clean = np.random.default_rng(0).normal(size=(2, 7, 5))
    sem_io.save_observations(str(tmp_path), tag, noisy, clean, 4e-4, 7)
    n2, c2, dt, nt = sem_io.load_observations(str(tmp_path), tag)
    np.testing.assert_array_equal(n2, noisy)
    assert dt == 4e-4 and nt == 7
    assert json.load(open(sem_io.observation_paths(str(tmp_path), tag)[2])) == {"obs_dt": 4e-4, "obs_nt": 7}
