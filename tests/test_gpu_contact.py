"""Static plane / box contact on the GPU (reference:
ChysX/src/collision/static_contact.{h,cu}, hooks cloth_simulator.cu:
997-1008, 1073-1078, 1189-1193): a cloth dropped onto a turned table box
over a ground plane, each substep started from the FP64 oracle's state,
against the oracle (relative max-norm position error <= 1e-5, PCG
iterations within +-1 of the FP32 oracle's); without and with Coulomb
friction.  Also the API surface the reference binds."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("friction", [0.0, 0.3])
def test_drape_substeps_match_oracle(gpu, friction):
    import torch

    from chysx import scenes

    spec = scenes.drape(friction=friction)
    dc = scenes.DeviceCloth(spec)
    o64 = helpers.oracle_from_spec(spec, precision=64)
    o32 = helpers.oracle_from_spec(spec, precision=32)
    worst, touching = 0.0, 0
    for s in range(90):
        po, vo = o64.state()
        dxo = o64.system()[3] if s else np.zeros((spec.n, 3))
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        o32.set_state(po, vo, o32.inv_mass())
        o32.set_dx(dxo)
        dc.step()
        o64.step(spec.dt)
        it32 = o32.step(spec.dt)
        p, _ = dc.state()
        ref = o64.state()[0]
        err = helpers.rel_maxnorm(p, ref)
        worst = max(worst, err)
        assert err <= 1e-5, (s, err)
        assert abs(dc.sim.last_pcg_iterations() - it32) <= 1, (s, dc.sim.last_pcg_iterations(), it32)
        touching = max(touching, int(np.sum(ref[:, 1] < 0.98 + 5e-3)))
    assert touching > 50, touching  # the cloth reached the table
    print(f"drape friction={friction}: worst one-substep rel err vs FP64 {worst:.3e}, {touching} vertices in contact")


def test_contact_api_surface(gpu):
    from chysx import scenes

    spec = scenes.drape()
    dc = scenes.DeviceCloth(spec)
    sim = dc.sim
    assert sim.static_plane_count() == 1 and sim.static_box_count() == 1
    assert sim.static_contact_thickness() == pytest.approx(5e-3)
    assert sim.static_contact_stiffness() == pytest.approx(1e4)
    assert sim.static_contact_friction() == 0.0
    assert sim.static_contact_friction_epsilon() == pytest.approx(1e-4)
    with pytest.raises(ValueError):
        sim.add_static_plane(np.zeros(2, np.float32), 0.0)
    sim.clear_static_shapes()
    assert sim.static_plane_count() == 0 and sim.static_box_count() == 0
    dc.step()  # no primitives: contact inactive, the step still runs
    # the cloth now falls freely: no vertex is pushed up
    p, v = dc.state()
    assert np.all(v[:, 1] <= 0.0)


def test_headless_export(gpu, tmp_path):
    """chysx.export writes each frame as a merged .bgeo (cloth + table box)
    and per-piece OBJ; the .bgeo's cloth points are the GPU state's bits."""
    from chysx import export, scenes
    from test_mesh_io import _attribute, _decode, _pairs

    spec = scenes.drape()
    paths = export.export_scene(spec, str(tmp_path), frames=3)
    assert len(paths) == 3 * 3
    g = _decode(str(tmp_path / "frame_0002.bgeo"))
    assert g["pointcount"] == spec.n + 8 and g["primitivecount"] == len(spec.tris) + 12
    P = _attribute(_pairs(g["attributes"])["pointattributes"][0])[2]["rawpagedata"].reshape(-1, 3)
    ref = scenes.DeviceCloth(spec)
    for _ in range(30):  # 3 frames of 1/60 s at dt = 1/600
        ref.step()
    assert np.array_equal(P[:spec.n], ref.state()[0].astype(np.float32))
    import chysx

    p_obj, t_obj = chysx.io.load_obj(str(tmp_path / "piece1_0002.obj"))
    assert p_obj.shape == (8, 3) and t_obj.shape == (12, 3)
