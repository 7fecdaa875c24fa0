"""The oracle against the reference's own known-answer solver tests
(src/test_ssol_solver*.c, src/test_ssol_by_receiver_integration.c), same
realisation counts and tolerances. This pins the oracle before it is used as
the GPU path's checker."""
import ref_checks as K


def test_solver1_all():
    K.solver1_all(10000, "oracle")


def test_solver1_primary():
    K.solver1_primary(10000, "oracle")


def test_solver1_atmosphere_zero():
    K.solver1_primary(10000, "oracle", variant="atm0")


def test_solver1_atmosphere():
    K.solver1_atm(10000, "oracle")


def test_solver1_monochromatic_spectral_extinction():
    K.solver1_mono(10000, "oracle")


def test_solver2_mesh_heliostats():
    K.solver2(10000, "oracle")


def test_solver2b_punched_heliostats():
    K.solver2b(50000, "oracle")


def test_solver3_punched_plane():
    K.solver3(20000, "oracle")


def test_solver5_parabolic_cylinder():
    K.solver5(10000, "oracle")


def test_solver8_carved_parabolic_cylinder():
    K.solver8(100000, "oracle")


def test_solver4_paraboloid_focus():
    K.solver4(100000, "oracle")


def test_solver6_shadowing_and_rng_chaining():
    K.solver6_chain(10000, "oracle")


def test_solver7_beam_down():
    K.solver7(10000, "oracle")


def test_solver9_cube():
    K.solver9(100000, "oracle")


def test_solver10_pillbox():
    K.solver10(10000, 100000, "oracle", "pillbox")


def test_solver11_buie():
    K.solver10(10000, 100000, "oracle", "buie")


def test_solver12_gaussian():
    K.solver10(10000, 100000, "oracle", "gaussian")


def test_by_receiver_integration():
    K.by_receiver(10000, "oracle")
