"""The reference's XML input format (io/reader.cpp, params.cpp) read into this package's case."""
import os
import textwrap

import numpy as np
import pytest

import pyoracle as po
from conftest import cases_module, package

inputs = __import__("importlib").import_module("3dnssolver_amd.inputs")

TUTORIAL = "/root/reference/tutorials/shearLayer/input.xml"

DOC = textwrap.dedent("""\
    <?xml version="1.0" ?>
    <Input>
        <Geometry> <Dim>3 </Dim> <Lx> 1.0 </Lx> <Ly>2.0</Ly> <Lz>	0.5	</Lz> </Geometry>
        <Grid> <Nx> 16 </Nx> <Ny> 12 </Ny> <Nz> 8 </Nz> </Grid>
        <FluidProperties> <Mu> 2E-4 </Mu> <Pr> 0.72 </Pr> <Gamma> 1.4 </Gamma> </FluidProperties>
        <Time> <EndT> 1.0 </EndT> <CFLmax> 0.5 </CFLmax> </Time>
        <OutputProperties> <FileFormat>	Binary	</FileFormat> <OutputInterval> 10 </OutputInterval> </OutputProperties>
        <InitialConditions> <Case> {case} </Case> </InitialConditions>
        <BoundaryConditions>
            <West> {x} </West> <East> {x} </East> <North> Neumann </North> <South> Neumann </South>
            <Front> Periodic </Front> <Back> Periodic </Back>
        </BoundaryConditions>
        <!-- <BoundaryConditions> <West> Inflow </West> </BoundaryConditions> -->
        <NumericalSchemes>
            <ConvectionScheme> <Scheme> {conv} </Scheme> <Order> {corder} </Order>
                <CharacteristicSpace> {char} </CharacteristicSpace> <WeightType> {w} </WeightType>
                <RiemannSolver> LaxFriedrichs </RiemannSolver> </ConvectionScheme>
            <DiffusionScheme> <Scheme> {diff} </Scheme> <Order> {dorder} </Order> </DiffusionScheme>
        </NumericalSchemes>
        <DomainDecomposition> <NprocX> 1 </NprocX> <NprocY> 1 </NprocY> <NprocZ> 1 </NprocZ> </DomainDecomposition>
    </Input>
    """)


def write(tmp_path, **kw):
    args = dict(case="case3", x="Periodic", conv="Weno", corder=5, char=1, w="Borges",
                diff="Centered", dorder=4)
    args.update(kw)
    p = os.path.join(str(tmp_path), "input.xml")
    open(p, "w").write(DOC.format(**args))
    return p


def test_reads_all_fields(tmp_path):
    case, extra = inputs.read_input_xml(write(tmp_path))
    assert (case["nx"], case["ny"], case["nz"], case["dim"]) == (16, 12, 8, 3)
    assert case["L"] == (1.0, 2.0, 0.5)
    assert (case["mu"], case["Pr"], case["gamma"], case["cfl"]) == (2e-4, 0.72, 1.4, 0.5)
    assert (case["conv"], case["weights"], case["diff_order"]) == ("weno5cw", "borges", 4)
    assert case["bc"] == ["periodic", "periodic", "neumann", "neumann", "periodic", "periodic"]
    assert extra == dict(end_t=1.0, io_interval=10, output_file_format="Binary", testcase="case3",
                         riemann_solver="LaxFriedrichs", nproc=(1, 1, 1))
    case, _ = inputs.read_input_xml(write(tmp_path, char=0, w="Shu", diff="none"))
    assert (case["conv"], case["weights"], case["diff_order"]) == ("weno5", "shu", 0)


@pytest.mark.parametrize("kw,err", [
    (dict(case="vortex"), "InputFileError"), (dict(conv="Upwind"), "InputFileError"),
    (dict(corder=3), "InputFileError"), (dict(w="Jiang"), "InputFileError"),
    (dict(diff="Spectral"), "InputFileError"), (dict(dorder=8), "InputFileError"),
    (dict(x="Wall"), "InputFileError"), (dict(conv="Centered", corder=4), "UnsupportedInput"),
    (dict(diff="Shen"), "UnsupportedInput"), (dict(x="Inflow"), "UnsupportedInput"),
])
def test_rejects_like_the_reference(tmp_path, kw, err):
    with pytest.raises(getattr(inputs, err)):
        inputs.read_input_xml(write(tmp_path, **kw))


def test_missing_element(tmp_path):
    p = write(tmp_path)
    s = open(p).read().replace("<Pr> 0.72 </Pr>", "")
    open(p, "w").write(s)
    with pytest.raises(inputs.InputFileError, match="Pr"):
        inputs.read_input_xml(p)


@pytest.mark.skipif(not os.path.exists(TUTORIAL), reason="reference tree not present")
def test_reads_the_reference_tutorial():
    case, extra = inputs.read_input_xml(TUTORIAL)
    assert (case["nx"], case["ny"], case["nz"], case["dim"]) == (100, 100, 1, 2)
    assert (case["mu"], case["Pr"], case["gamma"], case["cfl"]) == (1e-4, 0.71, 1.4, 0.4)
    assert (case["conv"], case["weights"], case["diff_order"]) == ("weno5", "shu", 6)
    assert case["bc"] == ["periodic"] * 6
    assert extra["testcase"] == "shear_xy" and extra["nproc"] == (2, 2, 1)


def test_thin_two_dimensional_block_on_the_oracle(tmp_path, oracle):
    # a 2-D input with Nz = 1, as the reference's 2-D tutorials are: the Riemann configuration
    # steps on the oracle and stays uniform in z
    C = cases_module()
    case, _ = inputs.read_input_xml(write(tmp_path))
    case = dict(case, dim=2, nz=1, L=(1.0, 1.0, 1.0 / 16), diff_order=4,
                bc=["neumann"] * 4 + ["periodic"] * 2)
    W = C.riemann_W(case, "case3")
    W1, U1, dts = oracle.step(case, W, 5)
    assert np.isfinite(po.interior(W1)).all()


def test_binary_output_matches_the_reference_writer(tmp_path):
    # the files of output.py are byte for byte those of the reference's BinaryWriter
    if not po.reference_available("strict"):
        pytest.skip("reference build (oracle/_ref) not present")
    out = __import__("importlib").import_module("3dnssolver_amd.output")
    C = cases_module()
    case = dict(C.base_case(6, 5, 4), L=(1.0, 1.3, 0.7))
    W = C.random_smooth_W(case, seed=3)
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    ref_dir.mkdir(), our_dir.mkdir()
    po.Reference("strict").write_binary(case, W, 7, ref_dir)
    out.write_mesh(case, str(our_dir))
    out.write_fields(W, str(our_dir), 7)
    for name in ("mesh_p_0.dat", "fields_7_p_0.dat"):
        assert (ref_dir / "results" / name).read_bytes() == (our_dir / "results" / name).read_bytes()
    np.testing.assert_array_equal(out.read_fields(str(our_dir / "results" / "fields_7_p_0.dat"), case),
                                  po.interior(W))


@pytest.mark.skipif(not os.path.exists(TUTORIAL), reason="reference tree absent")
def test_tutorial_document_is_the_reference_tutorial(tmp_path):
    # the GPU driver test's document holds the tutorial's values
    import importlib
    gd = importlib.import_module("test_gpu_driver")
    mine = inputs.read_input_xml(gd.write_xml(tmp_path))
    assert mine == inputs.read_input_xml(TUTORIAL)
