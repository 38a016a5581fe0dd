"""Copies the reference's benchmark input files (data, the drop-in input
format) into tests/golden/inputs and writes the integer wavenumber/mask golden
vectors.  Run in a container that has the reference checkout:

    python tests/golden/make_fixtures.py /root/reference

The wavenumber vectors restate the closed form of src/plane_layer/PL_geometry.f90:179-222
(k_index = 1..N; k = k_index-1 if k_index < NAA/2+1 else -(NAA+1-k_index);
dealiased iff |k*k_box| >= k_box*N/2), written out once so that the product's
integer indexing is compared with fixed vectors.
"""
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = {
    "heat_equation.coral.equations": "etc/benchmarks/00_heat_equation/coral.equations",
    "heat_equation.coral.parameters.in": "etc/benchmarks/00_heat_equation/coral.parameters.in",
    "stressfree_rb.coral.equations": "etc/benchmarks/01_Rayleigh_Benard_stressFree/coral.equations",
    "stressfree_rb.coral.parameters.in": "etc/benchmarks/01_Rayleigh_Benard_stressFree/coral.parameters.in",
    "rrbc_julien.coral.equations.velocityPressure": "etc/benchmarks/02_RRBC_julien_JFM96/coral.equations.velocityPressure",
    "rrbc_julien.coral.parameters.in": "etc/benchmarks/02_RRBC_julien_JFM96/coral.parameters.in",
    "rrbc_julien.coral.timeseries": "etc/benchmarks/02_RRBC_julien_JFM96/coral.timeseries",
    "rrbc_julien.coral.equations.toroidalPoloidal": "etc/benchmarks/02_RRBC_julien_JFM96/coral.equations.toroidalPoloidal",
    "rbc_vanilla.coral.equations": "etc/coral_equations_examples/coral.equations.RBC_primitiveVars_Vanilla",
    "example.coral.parameters.in": "etc/coral_equations_examples/coral.parameters.in",
    "rrbc_julien.coral.usrOutput": "etc/benchmarks/02_RRBC_julien_JFM96/coral.usrOutput",
    "example.coral.usrOutput": "etc/coral_equations_examples/coral.usrOutput",
    "stratified_channel.coral.equations.velocityPressure":
        "etc/benchmarks/coral.equations.stratifiedChannel.velocityPressure",
    "stratified_channel.coral.equations.toroidalPoloidal":
        "etc/benchmarks/coral.equations.stratifiedChannel.toroidalPoloidal",
    "baroclinic_turbulence.coral.equations": "etc/baroclinic_turbulence/coral.equations",
    "radiative_heating.coral.equations": "etc/coral_equations_examples/coral.equations.radiative_heating_potentials",
    "rbc_penalised.coral.equations": "etc/coral_equations_examples/coral.equations.RBC_penalised",
}


def wavenumbers(N):
    NAA = (3 * N) // 2
    out = []
    for k_index in range(1, NAA + 1):
        out.append(k_index - 1 if k_index < NAA // 2 + 1 else -(NAA + 1 - k_index))
    return out


def main(ref):
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    for dst, src in INPUTS.items():
        shutil.copyfile(os.path.join(ref, src), os.path.join(HERE, "inputs", dst))
    gold = {}
    for NX, NY in [(64, 64), (256, 256), (96, 64), (16, 32), (512, 512)]:
        kx = wavenumbers(NX)[: NX // 2]
        ky = wavenumbers(NY)
        gold[f"{NX}x{NY}"] = {
            "kx_int": kx,
            "ky_int": ky,
            "dealias_x": [int(2 * abs(k) >= NX) for k in kx],
            "dealias_y": [int(2 * abs(k) >= NY) for k in ky],
        }
    with open(os.path.join(HERE, "wavenumbers.json"), "w") as f:
        json.dump(gold, f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
