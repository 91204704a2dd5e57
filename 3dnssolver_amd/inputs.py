"""The reference's XML input file -> a case of this package.

Reads the document Reader::read_input_file parses (io/reader.cpp:3-172; example
tutorials/shearLayer/input.xml) with the same element names, whitespace collapsing of the texts
(tinyxml2 COLLAPSE_WHITESPACE) and the checks of Params::sanity_check (params.cpp:3-120), and maps
it onto the fields this package's Solver takes.  Choices the reference accepts but this path does
not implement (the Centered convection scheme, the Shen diffusion scheme, Inflow boundaries) raise
UnsupportedInput; invalid values raise InputFileError as the reference exits.  The rank counts
(DomainDecomposition) come back in `extra` and drive a decomposed run on one GPU (driver.py).
"""
import xml.etree.ElementTree as ET

VALID_TESTCASES = ("random", "sod_x", "sod_y", "sod_z", "case3", "case4", "case6", "case12",
                   "case15", "case17", "shear_xy", "jet", "rank", "shear_3d")


class InputFileError(ValueError):
    """Missing element or invalid value (the reference throws InputFileError or exits)."""


class UnsupportedInput(ValueError):
    """A valid reference input this path does not implement."""


def _text(parent, name):
    el = parent.find(name)
    if el is None:
        raise InputFileError(name)
    return " ".join((el.text or "").split())


def _int(parent, name):
    try:
        return int(_text(parent, name).split()[0])
    except (ValueError, IndexError):
        raise InputFileError(f"{name}: not an integer") from None


def _float(parent, name):
    try:
        return float(_text(parent, name).split()[0])
    except (ValueError, IndexError):
        raise InputFileError(f"{name}: not a number") from None


def _section(root, name):
    el = root.find(name)
    if el is None:
        raise InputFileError(name)
    return el


def read_input_xml(path):
    """Returns (case, extra): `case` the Solver fields (nx, ny, nz, dim, L, gamma, mu, Pr, cfl,
    conv, weights, diff_order, bc), `extra` the fields of the reference's Params this path does
    not consume (end_t, io_interval, output_file_format, testcase, riemann_solver, nproc)."""
    root = ET.parse(path).getroot()
    if root.tag != "Input":
        raise InputFileError("Input (first element)")
    geo, grid = _section(root, "Geometry"), _section(root, "Grid")
    fluid, time = _section(root, "FluidProperties"), _section(root, "Time")
    out, ic = _section(root, "OutputProperties"), _section(root, "InitialConditions")
    bcs, schemes = _section(root, "BoundaryConditions"), _section(root, "NumericalSchemes")
    dec = _section(root, "DomainDecomposition")

    conv_el = _section(schemes, "ConvectionScheme")
    conv_scheme = _text(conv_el, "Scheme")
    weights, riemann = "Shu", None
    if conv_scheme == "Weno":
        weights = _text(conv_el, "WeightType")
        riemann = _text(conv_el, "RiemannSolver")
    characteristic = _int(conv_el, "CharacteristicSpace")
    conv_order = _int(conv_el, "Order")
    diff_el = _section(schemes, "DiffusionScheme")
    diff_scheme = _text(diff_el, "Scheme")
    diff_order = _int(diff_el, "Order")
    faces = ("West", "East", "North", "South", "Front", "Back")
    bc = [_text(bcs, f) for f in faces]
    testcase = _text(ic, "Case")

    # Params::sanity_check (params.cpp:3-120)
    if testcase not in VALID_TESTCASES:
        raise InputFileError(f"Testcase: {testcase} is not valid")
    if conv_scheme not in ("Weno", "Centered", "none"):
        raise InputFileError(f"Convection scheme: {conv_scheme} is not valid")
    if conv_scheme == "Weno":
        if conv_order != 5:
            raise InputFileError(f"Convection scheme order: {conv_order} is not valid")
        if weights not in ("Shu", "Borges"):
            raise InputFileError(f"Weight type: {weights} is not valid")
    elif conv_order not in (2, 4, 6):
        raise InputFileError(f"Convection scheme order: {conv_order} is not valid")
    if diff_scheme not in ("Centered", "Shen", "none"):
        raise InputFileError(f"Diffusion scheme: {diff_scheme} is not valid")
    if diff_order not in (2, 4, 6):
        raise InputFileError(f"Diffusion scheme order: {diff_order} is not valid")
    for f, kind in zip(faces, bc):
        if kind not in ("Periodic", "Neumann", "Inflow"):
            raise InputFileError(f"Boundary condition {f}: {kind} is not valid")

    if conv_scheme == "Centered":
        raise UnsupportedInput("the Centered convection scheme is not on this path")
    if diff_scheme == "Shen":
        raise UnsupportedInput("the Shen diffusion scheme is not on this path")
    if "Inflow" in bc:
        raise UnsupportedInput("Inflow boundaries are not on this path")
    nproc = (_int(dec, "NprocX"), _int(dec, "NprocY"), _int(dec, "NprocZ"))
    case = dict(
        nx=_int(grid, "Nx"), ny=_int(grid, "Ny"), nz=_int(grid, "Nz"), dim=_int(geo, "Dim"),
        L=(_float(geo, "Lx"), _float(geo, "Ly"), _float(geo, "Lz")),
        gamma=_float(fluid, "Gamma"), mu=_float(fluid, "Mu"), Pr=_float(fluid, "Pr"),
        cfl=_float(time, "CFLmax"),
        conv="none" if conv_scheme == "none" else ("weno5cw" if characteristic == 1 else "weno5"),
        weights=weights.lower(),
        diff_order=0 if diff_scheme == "none" else diff_order,
        bc=[k.lower() for k in bc],
    )
    extra = dict(end_t=_float(time, "EndT"), io_interval=_int(out, "OutputInterval"),
                 output_file_format=_text(out, "FileFormat"), testcase=testcase,
                 riemann_solver=riemann, nproc=nproc)
    return case, extra
