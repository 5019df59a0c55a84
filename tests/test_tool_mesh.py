"""The friction-stir workload's inputs (CPU): the tool surface generator and
the plate lattice with the pin's volume left out.

Known answers: the mesh is a polygonal tool with n_theta sides, so its area
is exact in closed form (annuli n/2 sin(2 pi/n) (Ro^2 - Ri^2), the pin side
n 2 R sin(pi/n) L); every normal points out of the tool (-z on the shoulder
face and the pin tip, radially outward on the pin side), which is the side
the penalty force pushes particles to.
"""
import numpy as np

from weldformgpu_amd import scenarios


def _normals(nodes, el):
    a, b, c = nodes[el[:, 0]], nodes[el[:, 1]], nodes[el[:, 2]]
    n = np.cross(b - a, c - a)
    return (a + b + c) / 3.0, n, 0.5 * np.linalg.norm(n, axis=1)


def test_tool_mesh_area_and_normals():
    ctr, zs, Rs, Rp, L, nt = (0.01, 0.02), 0.006, 9e-3, 3e-3, 5e-3, 64
    nodes, el = scenarios.gen_tool_mesh(ctr, zs, Rs, Rp, L, n_theta=nt, n_rs=5, n_zp=7, n_rc=3)
    assert len(el) == 2 * nt * 5 + 2 * nt * 7 + nt + 2 * nt * 2
    cen, nrm, area = _normals(nodes, el)
    assert area.min() > 0.0
    poly = 0.5 * nt * np.sin(2 * np.pi / nt)
    exact = poly * (Rs ** 2 - Rp ** 2) + nt * 2 * Rp * np.sin(np.pi / nt) * L + poly * Rp ** 2
    assert abs(area.sum() - exact) < 1e-12 * exact
    u = nrm / np.linalg.norm(nrm, axis=1)[:, None]
    rad = cen[:, :2] - np.array(ctr)
    rad /= np.linalg.norm(rad, axis=1)[:, None] + 1e-300
    flat = np.isclose(np.ptp(nodes[el][:, :, 2], axis=1), 0.0)
    shoulder = flat & np.isclose(cen[:, 2], zs)
    tip = flat & np.isclose(cen[:, 2], zs - L)
    side = ~flat
    assert shoulder.sum() == 2 * nt * 5 and tip.sum() == nt + 2 * nt * 2
    assert side.sum() == 2 * nt * 7
    assert np.allclose(u[shoulder | tip], [0.0, 0.0, -1.0])
    assert np.all(np.abs(u[side][:, 2]) < 1e-12)
    assert np.all((u[side][:, :2] * rad[side]).sum(1) > np.cos(np.pi / nt) - 1e-9)


def test_friction_stir_plate_leaves_the_pin_empty():
    sc = scenarios.c5_friction_stir(scale=0.15, r=0.15e-3, n_theta=32, n_rs=4, n_zp=4, n_rc=2,
                                    thick=3.0)
    x = sc.fields["x"]
    cx, cy = sc.meta["tool_center"]
    rr = np.hypot(x[:, 0] - cx, x[:, 1] - cy)
    zb = sc.meta["shoulder_z"] - sc.meta["L_pin"]
    assert not np.any((rr < sc.meta["R_pin"]) & (x[:, 2] > zb + sc.meta["r"]))
    assert np.any(rr < sc.meta["R_pin"])  # the plate continues under the pin
    assert sc.n == len(x) == len(sc.fields["m"])
    # the shoulder face lies half a spacing above the top layer
    assert np.isclose(sc.meta["shoulder_z"] - x[:, 2].max(), sc.meta["r"])
    m = sc.meshes[0]
    assert m["omega"][2] > 100.0 and m["vel"][0] == 2e-3
