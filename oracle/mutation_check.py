"""Mutation check of the oracle's pins (test infrastructure).

Plants one plausible mistake at a time in a copy of lare_oracle.c, builds
the copy, and runs the CPU pin tests against it (LARE_ORACLE_LIB points the
test wrapper at the copy).  A mutant is "caught" when at least one pin
fails.  The oracle is only trusted where every mutant of this list is
caught; the output table is the record of that.

    python oracle/mutation_check.py            # all mutants
    python oracle/mutation_check.py M2 M6      # a subset

Each mutant names the reference lines whose restatement it breaks.
"""
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "lare_oracle.c")
CFLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-shared"]
PIN_TESTS = ["tests/test_pin_hydro.py", "tests/test_pin_transport.py", "tests/test_pin_oracle3d.py",
             "tests/test_pin_potential.py", "tests/test_pin_variants.py", "tests/test_pin_rke.py",
             "tests/test_pin_cauchy.py", "tests/test_pin_neutral.py", "tests/test_pin_open.py"]

# (name, what it breaks, [(old, new), ...]); each `old` occurs exactly once
MUTANTS = [
    ("M1", "vertex average of vy1 at the y face (xremap.f90:74-77): wrong y index",
     [("double vyb = (A(o->vy1, ix, iy, iz) + A(o->vy1, ixm, iy, iz) + A(o->vy1, ix, iy, izm) + A(o->vy1, ixm, iy, izm)) * 0.25;",
       "double vyb = (A(o->vy1, ix, iy, iz) + A(o->vy1, ixm, iy, iz) + A(o->vy1, ix, iy, izm) + A(o->vy1, ixm, iym, izm)) * 0.25;")]),
    ("M2", "12-edge heating: alpha1(ix,iyp,iz) for alpha1(ix,iy,izp) (lagran.F90:726, :920)",
     [("A(o->alpha1, ix, iy, izp) * a[4]", "A(o->alpha1, ix, iyp, iz) * a[4]")]),
    ("M3", "edge_viscosity: visc1 and visc2 swapped (lagran.F90:850-851, :547)",
     [("o->visc2_norm = 0.25 * (o->p.gamma + 1.0) * o->p.visc2;", "o->visc2_norm = 0.25 * (o->p.gamma + 1.0) * o->p.visc1;"),
      ("SQ(o->p.visc1 * cs_edge)", "SQ(o->p.visc2 * cs_edge)")]),
    ("M4", "temperature_bcs exchanges T instead of rho (boundary.f90:214, quirk Q3 'fixed')",
     [("for (int ax = 0; ax < 3; ++ax) exchange(o, o->rho, ax, 0);\n  mirror_cell(o, o->temperature);",
       "for (int ax = 0; ax < 3; ++ax) exchange(o, o->temperature, ax, 0);\n  mirror_cell(o, o->temperature);")]),
    ("M5", "x momentum-flux area dyb*dzc for dyc*dzc (xremap.f90:615)",
     [("double area = ax == 0 ? DYC(iy) * DZC(iz)", "double area = ax == 0 ? DYB(iy) * DZC(iz)")]),
    ("M6", "Larsen constant 4.285 for 42.85 (conduct.f90:90)",
     [("o->p.flux_limiter * 42.85", "o->p.flux_limiter * 4.285")]),
    ("M7", "y-face temperature gradient over two cells (conduct.f90:232)",
     [("tg = (A(T, ix, iyp, iz) - A(T, ix, iy, iz)) / DYC(iy);", "tg = (A(T, ix, iyp, iz) - A(T, ix, iym, iz)) / DYC(iy);")]),
    ("M8", "edge density: arithmetic for harmonic mean (lagran.F90:820-821)",
     [("double rho_edge = 2.0 * r1 * r2 / (r1 + r2);", "double rho_edge = 0.5 * (r1 + r2);")]),
    ("M9", "edge sound speed: max for min (lagran.F90:822)",
     [("double cs_edge = dmin(A(cs_v, i1, j1, k1), A(cs_v, i2, j2, k2));",
       "double cs_edge = dmax(A(cs_v, i1, j1, k1), A(cs_v, i2, j2, k2));")]),
    ("M10", "viscous force weight a4: alpha2(ix,iy,izp) for alpha2(ix,iyp,izp) (lagran.F90:763)",
     [("double a4 = A(o->alpha2, ix, iyp, izp) * DXC(ix) * DZC(iz);", "double a4 = A(o->alpha2, ix, iy, izp) * DXC(ix) * DZC(iz);")]),
    ("M11", "Larsen q_sp: gradient without the 1/2 of the centred difference (conduct.f90:93-95)",
     [("double q_spx = -k0 * temp * (A(T, ix + 1, iy, iz) - A(T, ix - 1, iy, iz)) * 0.5 / DXB(ix);",
       "double q_spx = -k0 * temp * (A(T, ix + 1, iy, iz) - A(T, ix - 1, iy, iz)) * 1.0 / DXB(ix);")]),
    ("M12", "RKL2 nu_j with b(j-1) for b(j-2) (conduct.f90:334)",
     [("nu[j] = -1.0 * ((dj - 1.0) / dj) * (b[j] / b[j - 2]);", "nu[j] = -1.0 * ((dj - 1.0) / dj) * (b[j] / b[j - 1]);")]),
    ("M13", "z-face transverse y gradient taken over one cell but divided by two (conduct.f90:274-277)",
     [("tb_m = (A(T, ix, iym, izp) + A(T, ix, iym, iz)) / 2.0;\n        tg_a2 = (tb_p - tb_m) / (DYC(iy) + DYC(iym));",
       "tb_m = (A(T, ix, iy, izp) + A(T, ix, iy, iz)) / 2.0;\n        tg_a2 = (tb_p - tb_m) / (DYC(iy) + DYC(iym));")]),
    ("M14", "z-face averaged bx with the x index unshifted (conduct.f90:258-259)",
     [("bxf = 0.25 * (A(o->bx, ix, iy, iz) + A(o->bx, ix, iy, izp) + A(o->bx, ixm, iy, iz) + A(o->bx, ixm, iy, izp));",
       "bxf = 0.25 * (A(o->bx, ix, iy, iz) + A(o->bx, ix, iy, izp) + A(o->bx, ixp, iy, iz) + A(o->bx, ixp, iy, izp));")]),
    ("M15", "energy flux donor density not upwinded (xremap.f90:594)",
     [("double rhou = AT(o->rho1, c) * r.vad_p + AT(o->rho1, p) * r.vad_m;", "double rhou = AT(o->rho1, c);")]),
    ("M16", "shock-viscosity threshold on the current cell width of the other axis (lagran.F90:825)",
     [("A(o->alpha1, ix, iy, iz) = edge_viscosity(o, cs_v, dvdots, DXB(ix), DXB(ix - 1),",
       "A(o->alpha1, ix, iy, iz) = edge_viscosity(o, cs_v, dvdots, DYB(iy), DXB(ix - 1),")]),
    ("M17", "potential_field: colouring in three dimensions (x1 restarted from z1 and iz, :141-143)",
     [("      for (int iz = 1; iz <= nz; ++iz) {\n        int x1 = z1;",
       "      for (int iz = 1; iz <= nz; ++iz) {\n        int x1 = (iz % 2) ? z1 : 3 - z1;")]),
    ("M18", "potential_field: flux patch subtracted (:285-288)",
     [("A(phi, ix, iy, 0) = A(phi, ix, iy, 1) + flux0[", "A(phi, ix, iy, 0) = A(phi, ix, iy, 1) - flux0[")]),
    ("M19", "potential_field: residual without the 1/2 (:152-155)",
     [("                       2.0 / (dx1 + dy1 + dz1);", "                       1.0 / (dx1 + dy1 + dz1);")]),
    ("M20", "potential_field: upper boundary zero-gradient instead of phi = 0 (:270-273)",
     [("        A(phi, ix, iy, nz + 1) = 0.0;", "        A(phi, ix, iy, nz + 1) = A(phi, ix, iy, nz);")]),
    ("M21", "SHOCKCOMPRESSION: MAX for MIN (lagran.F90:808)",
     [("if (o->p.shock_compression) dvdots = dmin(0.0, dvdots);", "if (o->p.shock_compression) dvdots = dmax(0.0, dvdots);")]),
    ("M22", "SHOCKLIMITER: rl from the node before the edge (lagran.F90:839)",
     [("rl = (dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx / (dxp * dv2);", "rl = (dvxm * dvx + dvyp * dvy + dvzp * dvz) * dx / (dxp * dv2);")]),
    ("M23", "boris: gamma_b from va_max^2 / w2 (lagran.F90:1107)",
     [("A(o->gamma_boris, ix, iy, iz) = 1.0 / (1.0 + w2 / va_max2);", "A(o->gamma_boris, ix, iy, iz) = 1.0 / (1.0 + va_max2 / w2);")]),
    ("M24", "boris: J x B unscaled (lagran.F90:375-377)",
     [("fx = fx + gb * (jy * bzv - jz * byv);", "fx = fx + 1.0 * (jy * bzv - jz * byv);")]),
    ("M25", "cooling_term: heating averaged with the old density (lagran.F90:468)",
     [("A(o->cool_term_v, ix, iy, iz) = aav * o->dt * A(o->visc_heat, ix, iy, iz) / A(o->rho, ix, iy, iz) +",
       "A(o->cool_term_v, ix, iy, iz) = aav * o->dt * A(o->visc_heat, ix, iy, iz) / A(o->rho, ix, iy, iz) / A(o->cv1, ix, iy, iz) * A(o->cv, ix, iy, iz) +")]),
    ("M26", "rke: ai with dm at the edge s for dm at s-1 (xremap.f90:676)",
     [("double ai = (AT(v, c) - AT(o->flux, m)) * AT(o->dm, m) / mc", "double ai = (AT(v, c) - AT(o->flux, m)) * AT(o->dm, c) / mc")]),
    ("M27", "rke: edge term added to the cells at sweep coordinate s instead of s+1 (xremap.f90:687-690)",
     [("            d[ax] += 1;\n", "            d[ax] += 0;\n")]),
    ("M28", "rke: transverse cells on the lower side of the edge (iym for iyp, xremap.f90:688)",
     [("d[t1] += q & 1;", "d[t1] -= q & 1;")]),
    ("M29", "rke: edges over the sweep range 0..n-2 instead of 0..n-1 (xremap.f90:665-670)",
     [("    hi[ax] -= 1;\n    const int t1", "    hi[ax] -= 2;\n    const int t1")]),
    ("M30", "energy_correction: the gains added as well as the losses (diagnostics.F90:616)",
     [("        if (d < 0.0) d = 0.0;\n", "")]),
    ("M31", "rke: node masses before the sweep for after it (xremap.f90:673-674)",
     [("double mc = AT(o->rho_v1, c) * AT(o->cv2, c);", "double mc = AT(o->rho_v, c) * AT(o->cv1, c);")]),
    ("M32", "CAUCHY: stretching term of By from dvx/dy instead of dvy/dx (lagran.F90:1037)",
     [("double w4 = b1x * dvydx + b1y * dvydy + b1z * dvydz;", "double w4 = b1x * dvxdy + b1y * dvydy + b1z * dvydz;")]),
    ("M33", "CAUCHY: By1 without the 1/(1 + div v dt/2) of compression (lagran.F90:1043)",
     [("A(o->by1, ix, iy, iz) = (b1y + w4 * o->dt2) / (1.0 + dv);", "A(o->by1, ix, iy, iz) = (b1y + w4 * o->dt2);")]),
    ("M34", "CAUCHY: jy with the sign of its first term flipped (lagran.F90:315)",
     [("double y = -(w2 - w1) / DXC(ix);", "double y = (w2 - w1) / DXC(ix);")]),
    ("M35", "CAUCHY: dvz/dx from a stencil one node off in x (lagran.F90:1013-1014)",
     [("vzbm = (A(vz, ixm, iy, iz) + A(vz, ixm, iym, iz) + A(vz, ixm, iy, izm) + A(vz, ixm, iym, izm)) * 0.25;",
       "vzbm = (A(vz, ix, iy, iz) + A(vz, ixm, iym, iz) + A(vz, ixm, iy, izm) + A(vz, ixm, iym, izm)) * 0.25;")]),
    ("M36", "FOURTHORDER: k3 weighted 1 instead of 2 (lagran.F90:1370)",
     [("2.0 * A(k[2][c], ix, iy, iz)", "1.0 * A(k[2][c], ix, iy, iz)")]),
    ("M37", "FOURTHORDER: the third stage from B1 + dt k2 instead of dt/2 (lagran.F90:1350)",
     [("const double h[3] = {half_dt, half_dt, dt};", "const double h[3] = {half_dt, dt, dt};")]),
    ("M38", "FOURTHORDER: the heating with dt instead of dt/6 (lagran.F90:1388)",
     [("A(c1, ixm, iy, izm) + A(c1, ix, iym, izm) + A(c1, ixm, iym, izm)) * dt6 /",
       "A(c1, ixm, iy, izm) + A(c1, ix, iym, izm) + A(c1, ixm, iym, izm)) * dt /")]),
    ("M39", "get_neutral: 0.5 for 0.25 in the exponent's radiation term (neutral.f90:152)",
     [("orc_exp((0.25 * (t_v / t_rad - 1.0) + 1.0) * p->t_bar / t_v)",
       "orc_exp((0.5 * (t_v / t_rad - 1.0) + 1.0) * p->t_bar / t_v)")]),
    ("M40", "get_neutral: the dilution factor left out (neutral.f90:148, :151)",
     [("const double t_rad = p->t_rad, dilution = 0.5;", "const double t_rad = p->t_rad, dilution = 1.0;")]),
    ("M41", "neutral_fraction: bisection stopped at a 1e-4 step (neutral.f90:186)",
     [("if (fabs(dx) < 1.0e-8 || fabs(fa) < 1.0e-8) break;", "if (fabs(dx) < 1.0e-4 || fabs(fa) < 1.0e-8) break;")]),
    ("M42", "neutral_fraction: the ionisation energy left out of the EOS residual (neutral.f90:183-184)",
     [("const double fa = x - gm1 * (e0 - (1.0 - xi_a) * ip) / (2.0 - xi_a);",
       "const double fa = x - gm1 * e0 / (2.0 - xi_a);")]),
    ("M43", "pressure of lagrangian_step without the ionisation energy (lagran.F90:93-94)",
     [("(A(o->energy, ix, iy, iz) - (1.0 - xi_at(o, IDX(ix, iy, iz))) * o->ip);",
       "(A(o->energy, ix, iy, iz) - (1.0 - xi_at(o, IDX(ix, iy, iz))) * 0.0);"),
      ("(e1 - (1.0 - xi_at(o, IDX(ix, iy, iz))) * o->ip) * gm1",
       "(e1 - (1.0 - xi_at(o, IDX(ix, iy, iz))) * 0.0) * gm1")]),
    ("M44", "conduction temperature with 1 - xi for 2 - xi (conduct.f90:72-73)",
     [("return (o->p.gamma - 1.0) / (2.0 - xi) * (e - (1.0 - xi) * o->ip);",
       "return (o->p.gamma - 1.0) / (1.0 - xi + 1.0e-3) * (e - (1.0 - xi) * o->ip);")]),
    ("M45", "no neutral_fraction between RKL2 stages (conduct.f90:404)",
     [("      if (o->pi_eos) neutral_fraction(o);\n      copy_all(o, Y[2], o->energy);",
       "      copy_all(o, Y[2], o->energy);")]),
    ("M46", "perpendicular_resistivity's vertex field 'corrected' to the bx and by faces (neutral.f90:97-98)",
     [("double bxv = A(o->bx, ix, iy, iz) + A(o->bx, ix, iyp, iz) + A(o->bz, ix, iy, izp) + A(o->bz, ix, iyp, izp);",
       "double bxv = A(o->bx, ix, iy, iz) + A(o->bx, ix, iyp, iz) + A(o->bx, ix, iy, izp) + A(o->bx, ix, iyp, izp);"),
      ("double byv = A(o->by, ix, iy, iz) + A(o->by, ixp, iy, iz) + A(o->bz, ix, iy, izp) + A(o->bz, ixp, iy, izp);",
       "double byv = A(o->by, ix, iy, iz) + A(o->by, ixp, iy, iz) + A(o->by, ix, iy, izp) + A(o->by, ixp, iy, izp);")]),
    ("M47", "rkstep (Cowling): eta_perp left out of the EMF (lagran.F90:1567-1572)",
     [("A(fxr, ix, iy, iz) = -((jpx + jqx) * e + jqx * ep) * DXC(ix) * 0.5;",
       "A(fxr, ix, iy, iz) = -((jpx + jqx) * e) * DXC(ix) * 0.5;"),
      ("A(fyr, ix, iy, iz) = -((jpy + jqy) * e + jqy * ep) * DYC(iy) * 0.5;",
       "A(fyr, ix, iy, iz) = -((jpy + jqy) * e) * DYC(iy) * 0.5;"),
      ("A(fzr, ix, iy, iz) = -((jpz + jqz) * e + jqz * ep) * DZC(iz) * 0.5;",
       "A(fzr, ix, iy, iz) = -((jpz + jqz) * e) * DZC(iz) * 0.5;")]),
    ("M48", "set_dt: the Cowling limit without eta_perp (lagran.F90:1132-1134)",
     [("dt3 = 0.2 * dxlocal / dmax(dmax(A(o->eta, ix, iy, iz), A(o->eta_perp, ix, iy, iz)), NONE_ZERO);",
       "dt3 = 0.2 * dxlocal / dmax(A(o->eta, ix, iy, iz), NONE_ZERO);")]),
    ("M49", "rkstep (Cowling): j_par normalised by |B| instead of |B|^2 (lagran.F90:1527-1529)",
     [("jpx = (jx * bxv + jy * byv + jz * bzv) * bxv / magn_b;", "jpx = (jx * bxv + jy * byv + jz * bzv) * bxv / sqrt(magn_b);"),
      ("jpy = (jx * bxv + jy * byv + jz * bzv) * byv / magn_b;", "jpy = (jx * bxv + jy * byv + jz * bzv) * byv / sqrt(magn_b);"),
      ("jpz = (jx * bxv + jy * byv + jz * bzv) * bzv / magn_b;", "jpz = (jx * bxv + jy * byv + jz * bzv) * bzv / sqrt(magn_b);")]),
    ("M50", "open_bcs: the far field read from the first ghost cell layer (openboundary.f90:50-59)",
     [("lfc = hi ? n + 2 : -1,", "lfc = hi ? n + 1 : 0,")]),
    ("M51", "open_bcs: the normal components not turned outwards on the low side (openboundary.f90:38-43)",
     [("const double sg = hi ? 1.0 : -1.0;", "const double sg = 1.0;")]),
    ("M52", "open_bcs_fast: the incoming characteristic's velocity from the interior (openboundary.f90:363-366)",
     [("      vstar[i] = b->uxfar;", "      vstar[i] = b->vxbc[1];")]),
    ("M53", "open_bcs_alfven: the incoming Alfven wave's invariant from the interior (openboundary.f90:430-431)",
     [("      lys[i] = lambdayfar;", "      lys[i] = -b->bybc[1] * b->bxbc[1];")]),
    ("M54", "open_bcs: the entropy wave's density jump over the interior sound speed (openboundary.f90:378)",
     [("const double rhog = rhostar[2] + (pg - (pstar[2] - pmagstar[2])) / (c0far * c0far);",
       "const double rhog = rhostar[2] + (pg - (pstar[2] - pmagstar[2])) / (c0 * c0);")]),
    ("M55", "open_bcs not called (lagran.F90:410-412)",
     [("  if (o->any_open) open_bcs(o);", "  if (0) open_bcs(o);")]),
]


def build(src_text, out_so):
    with tempfile.NamedTemporaryFile("w", suffix=".c", dir=HERE, delete=False) as fh:
        fh.write(src_text)
        cpath = fh.name
    try:
        subprocess.run(["gcc", *CFLAGS, "-I", HERE, "-o", out_so, cpath, "-lm"], check=True)
    finally:
        os.remove(cpath)


def run_pins(lib):
    env = dict(os.environ, LARE_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", *PIN_TESTS, "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        "--timeout", "300"], cwd=ROOT, env=env, capture_output=True, text=True)
    failed = [ln.split(" ")[1] for ln in r.stdout.splitlines() if ln.startswith("FAILED")]
    summary = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr.strip()[-200:]
    return r.returncode, failed, summary


def main(names):
    base = open(SRC).read()
    todo = [m for m in MUTANTS if not names or m[0] in names]
    rows = []
    with tempfile.TemporaryDirectory() as tmp:
        # the unmutated oracle must pass every pin
        so = os.path.join(tmp, "base.so")
        build(base, so)
        rc, failed, summary = run_pins(so)
        print(f"unmutated: {summary}", flush=True)
        if rc != 0:
            print("the unmutated oracle fails its pins:", failed)
            return 2
        for name, what, edits in todo:
            text = base
            for old, new in edits:
                if text.count(old) != 1:
                    print(f"{name}: pattern occurs {text.count(old)} times: {old[:60]}")
                    return 2
                text = text.replace(old, new)
            so = os.path.join(tmp, f"{name}.so")
            build(text, so)
            t0 = time.time()
            rc, failed, summary = run_pins(so)
            caught = rc != 0
            rows.append((name, what, caught, failed))
            print(f"{name:4s} {'caught' if caught else 'MISSED':6s} {time.time() - t0:5.1f}s  {what}", flush=True)
            for f in failed[:6]:
                print(f"       fails {f}", flush=True)
    missed = [r[0] for r in rows if not r[2]]
    print(f"{len(rows) - len(missed)}/{len(rows)} mutants caught" + (f"; missed: {' '.join(missed)}" if missed else ""))
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main(set(sys.argv[1:])))
