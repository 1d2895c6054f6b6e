"""Replay oracle: the reference's own move sequence, decision for decision.

The float-faithful restatement of MC_move (src/starrynight-montecarlo-core.c:
151-191) with the reference's MT19937 stream (seed 0xDEADBEEF + T) records
every attempted move: site, float proposal, float dE, the genrand_real2 value
when drawn, the decision.  sn_mc_replay re-evaluates each move in FP64 on the
GPU against the same lattice state.  Bar: no decision differs outside the float
rounding band of the reference's threshold (mismatches == 0), and the final
lattice equals the reference's float lattice exactly.
"""
import numpy as np
import pytest

from tests.helpers import print_case

pytestmark = pytest.mark.gpu

CASES = {
    # config: 2D 25x25x1, rc=3, J=1, E=0, T=300 (the reference's CPU-runnable case)
    "2d_25x25_T300": dict(X=25, Y=25, Z=1, rc=3, J=1.0, E=(0.0, 0.0, 0.0), K=0.0, T=300,
                          lengths=(1.0,), prev=(1.0,), cx=False, moves=20000),
    "3d_8cube_field_vacancy_K": dict(X=8, Y=8, Z=8, rc=3, J=1.0, E=(0.0, 0.0, 0.1), K=0.5, T=300,
                                     lengths=(1.0, 0.0), prev=(0.5, 0.5), cx=False, moves=20000),
    "3d_10cube_constrain_x": dict(X=10, Y=10, Z=10, rc=3, J=1.0, E=(0.02, 0.0, 0.0), K=0.0, T=150,
                                  lengths=(1.0,), prev=(1.0,), cx=True, moves=10000),
    "3d_12cube_rc4_T500": dict(X=12, Y=12, Z=12, rc=4, J=0.5, E=(0.0, 0.0, 0.0), K=0.0, T=500,
                               lengths=(1.0, 0.5), prev=(0.7, 0.3), cx=False, moves=10000),
}


@pytest.mark.parametrize("name", list(CASES))
def test_replay_reference_decisions(sn, oracle, name):
    c = CASES[name]
    print_case("replay", name=name, moves=c["moves"])
    T = c["T"]
    lat32, mt = oracle.random_lattice_f32(c["X"], c["Y"], c["Z"], 0xDEADBEEF + T, c["lengths"], c["prev"])
    start = lat32.copy()
    E32 = tuple(float(np.float32(v)) for v in c["E"])   # Efield is narrowed to float (config.c:132-134)
    o = oracle.Oracle(c["X"], c["Y"], c["Z"], c["rc"], J=c["J"], E=E32, K=c["K"], constrain_to_x=c["cx"],
                      beta=oracle.beta_of_T(T))
    res = o.mc_moves_f32(lat32, mt, c["moves"], trace=True)
    tr = res["trace"]
    live = tr[tr["vacancy"] == 0]
    rp = sn.ReplicaParams(c["X"], c["Y"], c["Z"], DipoleCutOff=c["rc"], CageStrain=c["J"], Efield=E32,
                          K=c["K"], ConstrainToX=c["cx"], T=T)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(start.astype(np.float64))
        dec, mism, ties = ctx.replay(live["site"], live["prop"].astype(np.float64), live["u"],
                                     live["u_drawn"], live["accepted"])
        final = ctx.get_lattice()
        cnt = ctx.counters()
    print(f"  moves {len(live)}  accepted {int(live['accepted'].sum())}  mismatches {mism}  near ties {ties}",
          flush=True)
    assert mism == 0
    assert ties <= max(3, len(live) // 2000)
    assert np.array_equal(dec, live["accepted"].astype(np.uint8))
    assert np.array_equal(final, lat32.astype(np.float64))
    assert cnt["accept"] == res["accept"] and cnt["reject"] == res["reject"]
