"""CPU checks of the host layer that need no device: a material of no cell
(the reference's MaterialLinearElastic3/4(name) constructors,
bind_py_material.cc:115,141) takes pixels and exposes its internal field
collection with the reference's names and values
(material_linear_elastic3.cc:57-68, material_linear_elastic4.cc:62-73,
bind_py_field_collection.cc)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def µ():
    import muspectre_amd
    return muspectre_amd


def _hooke(E, nu, d):
    """lambda I (x) I + 2 mu I4S, entry (i, j, k, l) at i + d j + d^2 k + d^3 l
    (Hooke's law for an isotropic solid)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    I = np.eye(d)
    C = (lam * np.einsum("ij,kl->ijkl", I, I)
         + mu * (np.einsum("ik,jl->ijkl", I, I) + np.einsum("il,jk->ijkl", I, I)))
    return C.transpose(3, 2, 1, 0).ravel()  # column-major (i fastest)


@pytest.mark.parametrize("d", [2, 3])
def test_le3_of_no_cell(µ, d):
    cls = getattr(µ.material, "MaterialLinearElastic3_{}d".format(d))
    mat = cls("free")
    assert isinstance(mat, getattr(µ.material, "Material_{}d".format(d)))
    assert mat.get_name() == "free" and mat.size() == 0
    params = [(210e9, .3), (70e9, .33), (100e9, .25)]
    for i, (E, nu) in enumerate(params):
        mat.add_pixel([i] + [0] * (d - 1), E, nu)
    assert mat.size() == 3
    col = mat.collection
    assert col.field_names == ["local stiffness tensor"]
    assert col.statefield_names == []
    f = col.get_real_field("local stiffness tensor")
    assert f.name == "local stiffness tensor" and f.nb_components == d ** 4 and f.size == 3
    arr = f.array
    assert arr.shape == (d ** 4, 3)
    for i, (E, nu) in enumerate(params):
        np.testing.assert_allclose(arr[:, i], _hooke(E, nu, d), rtol=1e-15, atol=0)
    with pytest.raises(RuntimeError):
        col.get_int_field("local stiffness tensor")
    with pytest.raises(RuntimeError):
        col.get_real_field("no such field")


@pytest.mark.parametrize("d", [2, 3])
def test_le4_of_no_cell(µ, d):
    cls = getattr(µ.material, "MaterialLinearElastic4_{}d".format(d))
    mat = cls("free4")
    E = np.array([210e9, 70e9])
    nu = np.array([.3, .45])
    for i in range(2):
        mat.add_pixel([0] * d, E[i], nu[i])
    col = mat.collection
    assert col.field_names == ["local first Lame constant",
                               "local second Lame constant(shear modulus)"]
    lam = col.get_real_field("local first Lame constant").array
    mu = col.get_real_field("local second Lame constant(shear modulus)").array
    np.testing.assert_array_equal(lam.ravel(), E * nu / ((1 + nu) * (1 - 2 * nu)))
    np.testing.assert_array_equal(mu.ravel(), E / (2 * (1 + nu)))


def test_le3_le4_pixels_need_moduli(µ):
    """As in the reference's binding, add_pixel of LE3 and LE4 takes the
    pixel's Young's modulus and Poisson ratio (bind_py_material.cc:121-127)."""
    for name in ("MaterialLinearElastic3_2d", "MaterialLinearElastic4_2d"):
        mat = getattr(µ.material, name)("x")
        with pytest.raises(TypeError):
            mat.add_pixel([0, 0])


def test_banner_prints_name_and_copyright(capsys):
    """bind_py_common.cc:158 binds common.hh's banner(name, year,
    copyright_holder): the notice names the program and the holder."""
    from muspectre_amd import _muSpectre
    _muSpectre.banner("demo", 2026, "someone")
    out = capsys.readouterr().out
    assert "µSpectre demo" in out and "Copyright © 2026  someone" in out
    _muSpectre.banner(name="demo", year=2026, copyright_holder="someone")
