"""Initial-state loading and warm starts (reference ``workflow.py:450-596``).

The reference's own cases (``tests/unit/test_inundation_initial_state.py``:
metadata entries such as ``saved_order`` and the benchmark fields are
ignored, dry cells lose their momentum), then every file format, ``wse`` and
``(u, v)`` inputs, the ``cell_id`` mapping, the original order, the depth
clamp and the Manning floor, and a warm start written by ``save_warm_start``
reloading to the same solver state.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pandas as pd
import pytest

from inundation_amd.workflow import WorkflowResult, load_initial_state, save_warm_start

ZB = np.asarray([0.0, 0.0], dtype=np.float32)
IDENT = np.asarray([0, 1], dtype=np.int32)


def _load(state, n=2, zb=ZB, order="solver", perm=IDENT, dry_tol=1e-4):  # noqa: ANN001, ANN202
    return load_initial_state(state, n, zb, order=order, id_col="cell_id", perm=perm, dry_tol=dry_tol)


def _base() -> dict[str, object]:
    return {
        "h": np.asarray([1.0, 0.0], dtype=np.float32),
        "hu": np.asarray([2.0, 3.0], dtype=np.float32),
        "hv": np.asarray([4.0, 5.0], dtype=np.float32),
    }


def test_saved_order_metadata_is_ignored() -> None:
    h, hu, hv, n = _load({**_base(), "saved_order": np.asarray("original")})
    assert h.tolist() == [1.0, 0.0] and hu.tolist() == [2.0, 0.0] and hv.tolist() == [4.0, 0.0] and n is None


def test_benchmark_metadata_is_ignored() -> None:
    meta = {
        "saved_order": np.asarray("solver"),
        "mesh_source": np.asarray("mesh.parquet"),
        "t_end_s": np.float64(259200),
        "dt_max": np.float64(0.25),
        "cfl_interval": np.int64(2000),
    }
    h, hu, hv, n = _load({**_base(), **meta})
    assert h.tolist() == [1.0, 0.0] and hu.tolist() == [2.0, 0.0] and hv.tolist() == [4.0, 0.0] and n is None


def test_unknown_fields_and_missing_fields_raise() -> None:
    with pytest.raises(ValueError, match="unsupported fields"):
        _load({**_base(), "velocity": np.zeros(2)})
    with pytest.raises(ValueError, match="'h' or 'wse'"):
        _load({"hu": np.zeros(2), "hv": np.zeros(2)})
    with pytest.raises(ValueError, match=r"\(hu,hv\) or \(u,v\)"):
        _load({"h": np.zeros(2), "hu": np.zeros(2)})
    with pytest.raises(ValueError, match="non-finite"):
        _load({**_base(), "h": np.asarray([np.nan, 1.0])})


def test_wse_and_velocities_and_manning() -> None:
    zb = np.asarray([1.0, 2.0], dtype=np.float32)
    state = {"wse": np.asarray([3.0, 1.5]), "u": np.asarray([0.5, 1.0]), "v": np.asarray([-1.0, 2.0]),
             "n_mann": np.asarray([0.03, 0.0])}
    h, hu, hv, n = _load(state, zb=zb)
    assert h.tolist() == [2.0, 0.0]  # below the bed clamps to dry
    assert hu.tolist() == [1.0, 0.0] and hv.tolist() == [-2.0, 0.0]
    assert n.tolist() == [pytest.approx(0.03), pytest.approx(1e-4)]


def test_cell_id_maps_rows_and_original_order_permutes() -> None:
    perm = np.asarray([2, 0, 1], dtype=np.int32)  # solver cell k is file cell perm[k]
    state = {"cell_id": np.asarray([2, 0, 1]), "h": np.asarray([30.0, 10.0, 20.0]),
             "hu": np.zeros(3), "hv": np.zeros(3)}
    h, *_ = _load(state, n=3, zb=np.zeros(3, np.float32), order="solver", perm=perm)
    assert h.tolist() == [10.0, 20.0, 30.0]
    h, *_ = _load(state, n=3, zb=np.zeros(3, np.float32), order="original", perm=perm)
    assert h.tolist() == [30.0, 10.0, 20.0]
    with pytest.raises(ValueError, match="cell_id"):
        _load({**state, "cell_id": np.asarray([0, 0, 1])}, n=3, zb=np.zeros(3, np.float32))
    with pytest.raises(ValueError, match="initial_state_order"):
        _load(_base(), order="hilbert")


@pytest.mark.parametrize("suffix", [".npz", ".npy", ".csv", ".parquet"])
def test_file_formats(tmp_path: Path, suffix: str) -> None:
    b = _base()
    p = tmp_path / f"state{suffix}"
    if suffix == ".npz":
        np.savez(p, **b, n_mann=np.asarray([0.05, 0.02]))
    elif suffix == ".npy":
        np.save(p, np.column_stack([b["h"], b["hu"], b["hv"], [0.05, 0.02]]))
    else:
        df = pd.DataFrame({**b, "n_mann": [0.05, 0.02]})
        df.to_csv(p, index=False) if suffix == ".csv" else df.to_parquet(p)
    h, hu, hv, n = _load(str(p))
    assert h.tolist() == [1.0, 0.0] and hu.tolist() == [2.0, 0.0] and hv.tolist() == [4.0, 0.0]
    assert n.tolist() == [pytest.approx(0.05), pytest.approx(0.02)]
    with pytest.raises(FileNotFoundError):
        _load(str(tmp_path / "absent.npz"))
    (tmp_path / "x.txt").write_text("")
    with pytest.raises(ValueError, match="Unsupported initial state format"):
        _load(str(tmp_path / "x.txt"))


def test_npy_shape_is_checked(tmp_path: Path) -> None:
    p = tmp_path / "s.npy"
    np.save(p, np.zeros((2, 5)))
    with pytest.raises(ValueError, match=r"\(N,3\)"):
        _load(str(p))


@pytest.mark.parametrize("order", ["original", "solver"])
def test_warm_start_reloads_to_the_same_solver_state(tmp_path: Path, order: str) -> None:
    rng = np.random.default_rng(2)
    n = 50
    perm = rng.permutation(n).astype(np.int32)
    h = rng.uniform(0.0, 2.0, n).astype(np.float32)
    h[::7] = 0.0
    hu = np.where(h > 1e-4, rng.normal(size=n), 0.0).astype(np.float32)
    hv = np.where(h > 1e-4, rng.normal(size=n), 0.0).astype(np.float32)
    res = WorkflowResult([], [], h, 0.0, 0.0, 0.0, 0, hu_final=hu, hv_final=hv)
    p = save_warm_start(tmp_path / "warm.npz", res, perm, order=order, mesh_source="m.parquet", t_end_s=3600.0)
    h2, hu2, hv2, n2 = load_initial_state(
        str(p), n, np.zeros(n, np.float32), order=order, id_col="cell_id", perm=perm, dry_tol=1e-4
    )
    assert np.array_equal(h2, h) and np.array_equal(hu2, hu) and np.array_equal(hv2, hv) and n2 is None
    with pytest.raises(ValueError, match="unknown warm-start metadata"):
        save_warm_start(tmp_path / "w2.npz", res, perm, colour="blue")
