"""The CSV sample files of the reference's SampleWriter.

Layout: the header of SampleWriter::openGroups / open_bayesW (samplewriter.cpp) and
one row per kept iteration, written as Eigen prints a row vector with
IOFormat(StreamPrecision, DontAlignCols, ", ", ", ") on a default std::ofstream
(writer.cpp:9-20). The value formatting is checked against a C++ program built
here that prints the same doubles through std::ostream; the field counts against
the reference's sample vector sizes (BayesRBase.cpp:165: 3M + 3 + G + F + N;
bayeswbase.cpp:758: 2M + 4 (+ F)); the cadence against iteration >= burn-in and
iteration % thin == 0 (BayesRBase.cpp:250).
"""
import shutil
import subprocess

import numpy as np
import pytest

import bayesrrcmd_amd as brr

SPECIALS = [0.0, -0.0, 1.0, 0.1, 1e-05, 1.5e-300, 123456789.0, 1234567.0, 123456.0, 0.000123456789,
            -2.5, 3.14159265358979, 1e21, float("inf"), float("-inf"), 7.0, 100.0, 0.999999949]


def cxx_format(values, tmp_path):
    """std::ostream's default formatting of each value (the reference's writer), one per line."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "fmt.cpp"
    src.write_text(
        "#include <iostream>\n#include <fstream>\n#include <cstdio>\n"
        "int main(int c, char** v){ std::ifstream in(v[1]); unsigned long long b;"
        " while (in >> b) { double d; static_assert(sizeof d == sizeof b, \"\"); __builtin_memcpy(&d, &b, 8);"
        " std::cout << d << '\\n'; } }\n")
    exe = tmp_path / "fmt"
    subprocess.run([gxx, "-O1", str(src), "-o", str(exe)], check=True)
    bits = tmp_path / "bits.txt"
    bits.write_text("\n".join(str(int(x)) for x in np.asarray(values, np.float64).view(np.uint64)))
    out = subprocess.run([str(exe), str(bits)], check=True, capture_output=True, text=True).stdout
    return out.split()


def test_bayesr_samples_layout_and_format(tmp_path):
    rng = np.random.default_rng(5)
    M, N, G = 7, 5, 2
    path = tmp_path / "r.csv"
    rows = []
    with brr.SampleWriter(path, "bayesr", M, N, G) as w:
        for it in range(12):
            if not brr.sample_due(it, 4, 3):
                continue
            vals = dict(mu=rng.normal(), beta=rng.normal(size=M) * 10.0 ** rng.integers(-8, 3, M),
                        sigmaE=rng.random(), sigmaG=rng.random(G), comp=rng.integers(0, 4, M),
                        acum=rng.random(M), eps=rng.normal(size=N))
            vals["beta"][:3] = SPECIALS[:3]
            w.write_bayesr(it, **vals)
            rows.append((it, vals))
    lines = path.read_text().splitlines()
    header = lines[0].split(",")
    assert len(header) == 3 * M + 3 + G + 0 + N
    assert header[:2] == ["iteration", "mu"] and header[2] == "beta[1]" and header[M + 2] == "sigmaE"
    assert header[M + 3: M + 3 + G] == ["sigmaG[1]", "sigmaG[2]"]
    assert header[M + 3 + G] == "comp[1]" and header[2 * M + 3 + G] == "acum[1]"
    assert header[-1] == f"epsilon[{N}]"
    assert [it for it, _ in rows] == [6, 9]          # burn-in 4, thin 3
    assert len(lines) == 1 + len(rows)
    flat = []
    for it, v in rows:
        flat += [it, v["mu"], *v["beta"], v["sigmaE"], *v["sigmaG"], *v["comp"].astype(float), *v["acum"], *v["eps"]]
    expect = cxx_format(flat, tmp_path)
    got = [f for line in lines[1:] for f in line.split(", ")]
    assert all(len(line.split(", ")) == len(header) for line in lines[1:])
    assert got == expect


def test_bayesw_samples_layout_and_format(tmp_path):
    rng = np.random.default_rng(6)
    M = 9
    path = tmp_path / "w.csv"
    flat = []
    with brr.SampleWriter(path, "bayesw", M) as w:
        for it in (0, 5):
            beta = np.array(SPECIALS[:M]) * rng.normal(size=M)
            comp = rng.integers(0, 4, M)
            a, mu, sb = 3.0 + rng.random(), 4.0 + rng.random(), rng.random() * 1e-4
            w.write_bayesw(it, a, mu, beta, comp, sb)
            flat += [it, a, mu, *beta, *comp.astype(float), sb]
    lines = path.read_text().splitlines()
    header = lines[0].split(",")
    assert len(header) == 2 * M + 4
    assert header[:3] == ["iteration", "alpha", "mu"] and header[-1] == "sigma_b"
    assert header[3 + M] == "comp[1]"
    got = [f for line in lines[1:] for f in line.split(", ")]
    assert got == cxx_format(flat, tmp_path)


def test_stream_format_specials(tmp_path):
    """Special values print as std::ostream prints them (6 significant digits, inf, -0)."""
    path = tmp_path / "s.csv"
    with brr.SampleWriter(path, "bayesw", len(SPECIALS)) as w:
        w.write_bayesw(0, 1.0, 2.0, np.array(SPECIALS), np.zeros(len(SPECIALS), np.int32), 0.5)
    fields = path.read_text().splitlines()[1].split(", ")[3: 3 + len(SPECIALS)]
    assert fields == cxx_format(SPECIALS, tmp_path)
    assert fields[:7] == ["0", "-0", "1", "0.1", "1e-05", "1.5e-300", "1.23457e+08"]


def test_bad_writer_calls(tmp_path):
    with pytest.raises(brr.BrrHipError):
        brr.SampleWriter(tmp_path / "x.csv", "bayesr", 3, 0, 1)        # BayesR needs N
    with pytest.raises(brr.BrrHipError):
        brr.SampleWriter(tmp_path / "no_such_dir" / "x.csv", "bayesw", 3)
    assert brr.lib().brr_samples_close(None) == brr._abi.EINVAL


@pytest.mark.parametrize("kind", ["bayesr", "bayesw"])
def test_binary_samples_hold_the_csv_rows_exactly(tmp_path, kind):
    """The binary file carries the same row values as the CSV, at full precision."""
    rng = np.random.default_rng(9)
    m, n, g, f = 7, 11, 2, 3
    paths = {b: tmp_path / f"s_{kind}_{int(b)}" for b in (False, True)}
    rows = []
    writers = {b: brr.SampleWriter(str(paths[b]), kind, m, n, g, f, binary=b) for b in (False, True)}
    for it in (0, 5, 10):
        beta = rng.normal(size=m)
        comp = rng.integers(0, 4, m).astype(np.int32)
        if kind == "bayesr":
            mu, se, sg, gm = rng.normal(), rng.random(), rng.random(g), rng.normal(size=f)
            acum, eps = rng.random(m), rng.normal(size=n)
            row = np.concatenate([[it, mu], beta, [se], sg, gm, comp, acum, eps])
            for w in writers.values():
                w.write_bayesr(it, mu, beta, se, sg, comp, acum, eps, gamma=gm)
        else:
            al, mu, sb, th = rng.random() + 1, rng.normal(), rng.random(), rng.normal(size=f)
            row = np.concatenate([[it, al, mu], th, beta, comp, [sb]])
            for w in writers.values():
                w.write_bayesw(it, al, mu, beta, comp, sb, theta=th)
        rows.append(row)
    for w in writers.values():
        w.close()
    head, got = brr.read_samples_binary(str(paths[True]))
    want = np.array(rows)
    assert head == dict(kind=2 if kind == "bayesr" else 3, M=m, N=n, G=g, F=f)
    assert got.shape == want.shape == (3, (3 + 3 * m + g + f + n) if kind == "bayesr" else (4 + 2 * m + f))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    csv = np.array([[float(x) for x in line.split(", ")] for line in paths[False].read_text().splitlines()[1:]])
    assert np.allclose(csv, want, rtol=1e-5, atol=0)          # the CSV keeps 6 significant digits
    with pytest.raises(brr.BrrHipError):
        brr.read_samples_binary(str(paths[False]))
