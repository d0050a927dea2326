"""CPU checks of the file formats of the reference's workflow (PDB water box, OpenMM force-field
XML, the experimental RDF table) and of the oracle's Monte Carlo barostat
(water/water.jl:789-865). Files of the reference are read where they are present (this container);
the round trips need nothing from it."""
import math
import os

import numpy as np
import pytest

REF = "/root/reference/water"


def test_pdb_round_trip(tmp_path):
    import revsim
    from revsim import io

    wb = revsim.water_box(216, 1.862)
    p = str(tmp_path / "w.pdb")
    io.write_pdb(p, wb)
    back = io.read_pdb(p)
    assert back.n_molecules == 216 and np.allclose(back.box, wb.box, atol=5e-5)
    d = back.coords - np.mod(wb.coords.reshape(-1, 3), wb.box).reshape(-1)
    d -= 1.862 * np.round(d / 1.862)
    assert np.abs(d).max() <= 5.01e-5  # three decimals in Å
    assert np.array_equal(back.velocities, np.zeros_like(back.coords))


def test_pdb_rejects_wrong_atom_order(tmp_path):
    from revsim import io

    p = tmp_path / "bad.pdb"
    p.write_text("CRYST1   30.000   30.000   30.000  90.00  90.00  90.00 P 1           1 \n"
                 "ATOM      1  H1  HOH A   1       4.025  14.428  14.348  1.00  0.00\n"
                 "ATOM      2  O   HOH A   1       4.125  13.679  13.761  1.00  0.00\n"
                 "ATOM      3  H2  HOH A   1       4.670  13.062  14.249  1.00  0.00\n")
    with pytest.raises(ValueError):
        io.read_pdb(str(p))


@pytest.mark.skipif(not os.path.exists(f"{REF}/tip3p.pdb"), reason="reference files absent")
def test_reference_pdb():
    """water/tip3p.pdb: 895 TIP3P waters in a 3 nm box (water/water.jl:26-34)."""
    from revsim import io

    wb = io.read_pdb(f"{REF}/tip3p.pdb", temperature=295.15)
    assert wb.n_molecules == 895 and np.allclose(wb.box, 3.0)
    r, a = io.oh_geometry(wb)
    assert np.all(np.abs(r - 0.09572) < 5e-4)
    assert np.all(np.abs(np.degrees(a) - 104.52) < 0.2)
    # Maxwell-Boltzmann at 295.15 K: kinetic energy per degree of freedom ≈ kT/2
    m = wb.masses()
    ke = 0.5 * np.sum(m[:, None] * wb.velocities.reshape(-1, 3) ** 2) / (3 * wb.n_atoms)
    assert ke == pytest.approx(0.5 * 0.00831446262 * 295.15, rel=0.05)


def test_forcefield_xml_round_trip(tmp_path):
    import revsim
    from revsim import io

    th = revsim.DEFAULT_THETA * np.array([1.01, 0.97, 1.03, 0.9, 1.02, 1.1])
    p = str(tmp_path / "ff.xml")
    io.write_forcefield_xml(p, th, theta0=1.9)
    th2, t0 = io.read_forcefield_xml(p)
    assert np.array_equal(th2, th) and t0 == 1.9


@pytest.mark.skipif(not os.path.exists(f"{REF}/trained_force_fields/lj.xml"), reason="reference files absent")
def test_reference_forcefield_xml(tmp_path):
    """The reference's starting (tip3p.xml) and trained (lj.xml) force fields: θ read back, and
    writing them reproduces the files' values and layout."""
    import xml.etree.ElementTree as ET

    from revsim import io

    th, t0 = io.read_forcefield_xml(f"{REF}/trained_force_fields/tip3p.xml")
    assert np.allclose(th, [0.315075240503, 0.635968, 0.417, 462750.4, 0.09572, 836.8], rtol=1e-9)
    assert t0 == pytest.approx(1.82421813418, rel=1e-12)
    th, t0 = io.read_forcefield_xml(f"{REF}/trained_force_fields/lj.xml")
    assert th[2] == pytest.approx(0.45603299169668643, rel=1e-15)
    p = str(tmp_path / "lj.xml")
    io.write_forcefield_xml(p, th, t0)
    ours, ref = ET.parse(p).getroot(), ET.parse(f"{REF}/trained_force_fields/lj.xml").getroot()
    assert [e.tag for e in ours.iter()] == [e.tag for e in ref.iter()]
    for a, b in zip(ours.iter(), ref.iter()):
        for k, v in b.attrib.items():
            try:
                assert float(a.attrib[k]) == pytest.approx(float(v), rel=1e-15), (b.tag, k)
            except ValueError:
                assert a.attrib[k] == v, (b.tag, k)


@pytest.mark.skipif(not os.path.exists(f"{REF}/soper_2013_water_rdf.txt"), reason="reference files absent")
def test_reference_rdf_table():
    """read_rdf_data (water/water.jl:616-629): g_OO[86:334] holds 248 values on 0.261-1.002 nm;
    the O-O first peak of liquid water sits near 0.28 nm with g ≈ 2.5-3."""
    from revsim import io

    r, oo, oh, hh = io.read_rdf_data(f"{REF}/soper_2013_water_rdf.txt")
    assert r[86] == pytest.approx(0.261) and r[333] == pytest.approx(1.002) and oo[86:334].size == 248
    g = io.rdf_on_bins(r, oo, 200, 0.8)
    xb = (np.arange(200) + 0.5) * 0.004
    k = int(np.argmax(g))
    assert 0.27 < xb[k] < 0.29 and 2.3 < g[k] < 3.2
    assert np.all(g[xb < 0.22] < 0.05)


def _baro_orc(P, every=10, n=216, box=1.862, rc=0.8, **kw):
    import oracle as O
    import revsim

    o = O.Oracle(revsim.make_config(n, box, rc, barostat_every=every, barostat_pressure=P, **kw))
    o.set_theta(revsim.DEFAULT_THETA)
    return o


def test_barostat_follows_the_pressure(box216):
    """Large positive pressure compresses, large negative pressure expands; at ~1 bar the 216-water
    lattice start relaxes a little. The molecules stay whole (bond lengths unchanged by moves)."""
    import revsim
    from revsim import io

    V0 = 1.862 ** 3
    Vs = {}
    for P in (-5e4, 5e4):
        o = _baro_orc(P)
        x, v, _ = o.forward(box216.coords, box216.velocities, 13, 100)
        b = o.get_box()
        Vs[P] = float(np.prod(b))
        r, _ = io.oh_geometry(revsim.WaterBox(x, v, b))
        assert np.all(np.abs(r - 0.0957) < 0.01)
    assert Vs[5e4] < V0 < Vs[-5e4]


def test_barostat_trajectory_has_no_reverse(box216):
    import oracle as O

    o = _baro_orc(1.0)
    xN, vN, _ = o.forward(box216.coords, box216.velocities, 13, 20)
    with pytest.raises(O.OracleError):
        o.reverse(xN, vN, 13, 20)
