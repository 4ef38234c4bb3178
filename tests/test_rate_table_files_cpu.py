"""Rate-table files on the host (no GPU): the layout of scattering_struct::save
(src/monte_carlo/scattering_struct.h:55-94) written by kmc_save_rate_table and read back by
kmc_load_rate_table.  Pinned by an exact round trip, by the reference's own reader
(visualization/monte_carlo_results.py:263-282, restated below with numpy) applied to our files,
and by a hand-written file in the format the reference itself emits (Armadillo's 4-digit vector
print for the axes, the stream's default 6 significant digits for the rates)."""
import numpy as np
import pytest

NM = 1e-9


def _table(rng, shape=(3, 4, 2, 5)):
    axes = [np.sort(rng.uniform(-10, 10, n)) * (NM if k else 1.0) for k, n in enumerate(shape)]
    rates = 10.0 ** rng.uniform(-300, 15, shape)
    rates.flat[::7] = 0.0                      # zeros survive
    rates.flat[3] = 5e-324                     # and the smallest subnormal
    return axes, rates


def test_round_trip_is_exact(kmc, tmp_path):
    rng = np.random.default_rng(4)
    axes, rates = _table(rng)
    kmc.save_rate_table(tmp_path, *axes, rates)
    got = kmc.load_rate_table(tmp_path)
    for a, b in zip(got[:4], axes):
        assert np.array_equal(a, b)
    assert got[4].shape == rates.shape and np.array_equal(got[4], rates)
    # a second table under another name in the same directory
    kmc.save_rate_table(tmp_path, *axes, rates * 2, name="scat_table_1_2")
    assert np.array_equal(kmc.load_rate_table(tmp_path, "scat_table_1_2")[4], rates * 2)
    assert np.array_equal(kmc.load_rate_table(tmp_path)[4], rates)


def test_reference_reader_reads_our_files(kmc, tmp_path):
    # monte_carlo_results.py:263-282: loadtxt(skiprows=4), dims from the third line,
    # reshape in C order; the axes by loadtxt
    rng = np.random.default_rng(5)
    axes, rates = _table(rng, (21, 3, 2, 4))
    kmc.save_rate_table(tmp_path, *axes, rates)
    f = tmp_path / "scat_table.rates.dat"
    lines = f.read_text().splitlines()
    assert lines[0] == "sizes:" and lines[1] == "theta, z_shift, axis_shift_1, axis_shift_2"
    assert lines[3] == ""
    dims = [int(d) for d in lines[2].split(",")]
    assert dims == [21, 3, 2, 4]
    r = np.loadtxt(f, skiprows=4).reshape(dims, order="C")
    assert np.array_equal(r, rates)
    for name, a in zip(("theta", "z_shift", "axis_shift_1", "axis_shift_2"), axes):
        assert np.array_equal(np.atleast_1d(np.loadtxt(tmp_path / f"scat_table.{name}.dat")), a)


def test_reads_the_reference_output_format(kmc, tmp_path):
    # what scattering_struct::save writes: `file << vec` (Armadillo: one element per line,
    # right-aligned, 4 decimals or scientific) and `file << rate << "\n"` (6 significant digits)
    (tmp_path / "scat_table.theta.dat").write_text("        0\n   1.5708\n   3.1416\n")
    (tmp_path / "scat_table.z_shift.dat").write_text("   1.5000e-09\n   1.0000e-08\n")
    (tmp_path / "scat_table.axis_shift_1.dat").write_text("  -1.0000e-08\n")
    (tmp_path / "scat_table.axis_shift_2.dat").write_text("  -1.0000e-08\n            0\n")
    vals = [1.23457e+12, 0, 9.87654e-05, 42, 3.14159e+13, 2.71828, 1e-300, 5.5e+14, 0, 1, 2, 3]
    (tmp_path / "scat_table.rates.dat").write_text(
        "sizes:\ntheta, z_shift, axis_shift_1, axis_shift_2\n3,2,1,2\n\n" +
        "".join(f"{v:g}\n" for v in vals))
    th, z, a1, a2, r = kmc.load_rate_table(tmp_path)
    assert np.array_equal(th, [0.0, 1.5708, 3.1416])
    assert np.array_equal(z, [1.5e-9, 1e-8]) and np.array_equal(a1, [-1e-8])
    assert np.array_equal(a2, [-1e-8, 0.0])
    assert r.shape == (3, 2, 1, 2)
    assert np.array_equal(r.ravel(), [float(f"{v:g}") for v in vals])
    assert r[1, 0, 0, 1] == 2.71828 and r[2, 1, 0, 1] == 3.0


def test_errors(kmc, tmp_path):
    with pytest.raises(kmc.KmcError, match="cannot open"):
        kmc.load_rate_table(tmp_path)
    rng = np.random.default_rng(6)
    axes, rates = _table(rng, (2, 2, 2, 2))
    kmc.save_rate_table(tmp_path, *axes, rates)
    f = tmp_path / "scat_table.rates.dat"
    good = f.read_text()
    f.write_text(good.replace("sizes:", "size"))
    with pytest.raises(kmc.KmcError, match="bad header"):
        kmc.load_rate_table(tmp_path)
    f.write_text(good.rsplit("\n", 2)[0] + "\n")        # one rate short
    with pytest.raises(kmc.KmcError, match="holds 15 rates"):
        kmc.load_rate_table(tmp_path)
    f.write_text(good + "1.0\n")                         # one rate too many
    with pytest.raises(kmc.KmcError, match="holds 17 rates"):
        kmc.load_rate_table(tmp_path)
    f.write_text(good.rsplit("\n", 2)[0] + "\nabc\n")  # the last rate unreadable
    with pytest.raises(kmc.KmcError, match="bad number"):
        kmc.load_rate_table(tmp_path)
    f.write_text(good)
    (tmp_path / "scat_table.axis_shift_1.dat").write_text("0\n1\n2\n")
    with pytest.raises(kmc.KmcError, match="holds 3 values, the header gives 2"):
        kmc.load_rate_table(tmp_path)
    with pytest.raises(ValueError, match="does not match"):
        kmc.save_rate_table(tmp_path, *axes, rates[:, :1])
    with pytest.raises(kmc.KmcError, match="cannot write"):
        kmc.save_rate_table(tmp_path / "missing_dir", *axes, rates)


def test_load_with_wrong_dims_is_a_capacity_error(kmc, tmp_path):
    import ctypes as C
    rng = np.random.default_rng(7)
    axes, rates = _table(rng, (2, 3, 2, 2))
    kmc.save_rate_table(tmp_path, *axes, rates)
    lib = kmc.library()
    dims = np.array([2, 3, 2, 3], np.int64)               # n_a2 too large
    out = [np.zeros(8) for _ in range(4)] + [np.zeros(64)]
    st = lib.kmc_load_rate_table(str(tmp_path).encode(), None,
                                 dims.ctypes.data_as(C.POINTER(C.c_int64)),
                                 *[o.ctypes.data_as(C.POINTER(C.c_double)) for o in out])
    assert st == kmc.ERR_CAPACITY
    assert list(dims) == [2, 3, 2, 3]                      # left as given
