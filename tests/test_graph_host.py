"""The library's graph validation, derived properties and the Python graph model, on the CPU.

syf_graph_info validates and compiles without a device.  The refusals restate the reference's
validation tests (validate.rs:497-771: the English messages of the reference's catalogue name
the same faults and nodes); the derived properties (step scale, iteration product, trap scales)
are checked against the oracle's restatement of eval.rs / validate.rs on every bundled graph and
on the graphs of graph_parity.rs.
"""
import os
import re

import pytest

from oracle import oracle
from syfrack_amd.graph import Graph, Node

from test_oracle_graph import SPHERE, mk

REF = "/root/reference"


def info_error(g):
    with pytest.raises(ValueError) as e:
        g.info()
    return str(e.value)


def node_of(msg):
    m = re.match(r"graph: node (\d+): ", msg)
    return int(m.group(1)) if m else None


def minimal():
    return Graph.new_minimal()


def test_the_minimal_graph_is_valid():
    step, prod, scales = minimal().info()
    assert step == 1.0 and prod == 1 and scales == (1.0, 1.0, 1.0, 1.0)


def test_a_cycle_is_refused():
    g = minimal()
    g.nodes += [Node(3, "BoxFold", {"limit": 1.0}), Node(4, "Scale", {"factor": 2.0})]
    g.links = [((0, 0), (3, 0)), ((3, 0), (4, 0)), ((4, 0), (3, 0)), ((4, 0), (1, 0)), ((1, 0), (2, 0))]
    assert "Cycle" in info_error(g)


def test_a_mistyped_link_is_refused():
    g = minimal()
    g.links = [((0, 0), (2, 0)), ((1, 0), (2, 0))]
    msg = info_error(g)
    assert "type" in msg and node_of(msg) == 2


def test_two_links_into_one_port_are_refused():
    g = minimal()
    g.nodes.append(Node(3, "Sphere", {"radius": 2.0}))
    g.links += [((0, 0), (3, 0)), ((3, 0), (2, 0))]
    assert "port" in info_error(g)


def test_missing_or_duplicate_output_is_refused():
    g = minimal()
    g.nodes = [n for n in g.nodes if n.kind != "Output"]
    g.links = [l for l in g.links if l[1][0] != 2]
    assert "Output" in info_error(g)
    g = minimal()
    g.nodes.append(Node(3, "Output"))
    assert "Output" in info_error(g)


def test_an_iterate_body_only_holds_point_to_point_nodes():
    g = minimal()
    g.nodes += [Node(3, "Iterate", {"iterations": 4, "bailout": 4.0, "body": [4]}), Node(4, "Sphere", {"radius": 1.0})]
    g.links = [((0, 0), (3, 0)), ((3, 0), (1, 0)), ((1, 0), (2, 0))]
    msg = info_error(g)
    assert "Repeat" in msg and node_of(msg) == 4


def test_body_nodes_cannot_also_be_wired():
    g = minimal()
    g.nodes += [Node(3, "Iterate", {"iterations": 4, "bailout": 4.0, "body": [4]}), Node(4, "BoxFold", {"limit": 1.0})]
    g.links = [((0, 0), (3, 0)), ((3, 0), (1, 0)), ((1, 0), (2, 0)), ((0, 0), (4, 0))]
    assert node_of(info_error(g)) == 4


def test_limits_are_enforced():
    g = minimal()
    g.nodes += [Node(3, "Iterate", {"iterations": 16, "bailout": 4.0, "body": []}),
                Node(4, "Iterate", {"iterations": 32, "bailout": 4.0, "body": [3]})]
    g.links = [((0, 0), (4, 0)), ((4, 0), (1, 0)), ((1, 0), (2, 0))]
    assert "Iteration" in info_error(g)  # 16 x 32 = 512 > 256
    g = minimal()
    g.nodes[1] = Node(1, "Scale", {"factor": 0.0})
    assert node_of(info_error(g)) == 1


def test_a_dangling_node_is_inactive_but_valid():
    g = minimal()
    g.nodes.append(Node(3, "BoxFold", {"limit": 1.0}))
    assert g.info()[1] == 1


def test_out_of_bounds_parameter_is_reported_with_its_path():
    g = minimal()
    g.nodes[1].params["radius"] = 1000.0
    msg = info_error(g)
    assert "nodes[1].radius" in msg and node_of(msg) == 1
    assert msg == "graph: node 1: parameter `fractal.nodes[1].radius` = 1000 is outside [0.01, 100]"
    g = minimal()
    g.safety = 1.5
    assert "fractal.safety" in info_error(g)


def test_a_dangling_iterate_does_not_count_in_the_iteration_product():
    g = minimal()
    g.nodes.append(Node(3, "Iterate", {"iterations": 64, "bailout": 4.0, "body": []}))
    assert g.info()[1] == 1
    g.nodes.append(Node(4, "Iterate", {"iterations": 64, "bailout": 4.0, "body": [3]}))
    assert g.info()[1] == 1


def test_a_node_in_two_bodies_or_twice_in_one_is_refused():
    g = minimal()
    g.nodes += [Node(3, "Iterate", {"iterations": 2, "bailout": 4.0, "body": [4]}), Node(4, "Scale", {"factor": 2.0}),
                Node(5, "Iterate", {"iterations": 2, "bailout": 4.0, "body": [4]})]
    g.links = [((0, 0), (3, 0)), ((3, 0), (5, 0)), ((5, 0), (1, 0)), ((1, 0), (2, 0))]
    msg = info_error(g)
    assert "two bodies" in msg and node_of(msg) == 4
    g = minimal()
    g.nodes += [Node(3, "Iterate", {"iterations": 2, "bailout": 4.0, "body": [4, 4]}), Node(4, "Scale", {"factor": 2.0})]
    g.links = [((0, 0), (3, 0)), ((3, 0), (1, 0)), ((1, 0), (2, 0))]
    msg = info_error(g)
    assert "twice" in msg and node_of(msg) == 4
    g.node(3).params["body"] = [4]
    g.info()


def test_an_incomplete_graph_is_refused_naming_the_unwired_node():
    g = minimal()
    g.links = [l for l in g.links if l[1][0] != 2]
    assert node_of(info_error(g)) == 2
    g = minimal()
    g.nodes.append(Node(3, "Union"))
    g.links = [l for l in g.links if l[1][0] != 2] + [((1, 0), (3, 0)), ((3, 0), (2, 0))]
    assert node_of(info_error(g)) == 3


def test_a_dangling_cycle_is_refused():
    g = minimal()
    g.nodes += [Node(3, "BoxFold", {"limit": 1.0}), Node(4, "Scale", {"factor": 2.0})]
    g.links += [((3, 0), (4, 0)), ((4, 0), (3, 0))]
    assert "Cycle" in info_error(g)


def test_node_faults_name_their_node():
    g = minimal()
    g.nodes.append(Node(3, "Scale", {"factor": 0.0}))
    assert node_of(info_error(g)) == 3
    g = minimal()
    g.nodes += [Node(3, "Iterate", {"iterations": 4, "bailout": 4.0, "body": [4]}), Node(4, "Sphere", {"radius": 1.0})]
    assert node_of(info_error(g)) == 4


def test_active_instances_are_refused_by_name():
    g = minimal()
    g.nodes.append(Node(3, "Instances", {"half": (1.0, 1.0, 1.0), "margin": 0.1}))
    g.links = [((0, 0), (1, 0)), ((1, 0), (3, 0)), ((3, 0), (2, 0))]
    msg = info_error(g)
    assert node_of(msg) == 3 and "Instances" in msg
    g.node(3).enabled = False  # a disabled Instances passes its input through
    assert g.info()[1] == 1


GRAPHS = Graph.bundled_names()


@pytest.mark.parametrize("name", GRAPHS)
def test_derived_properties_match_the_oracle(name):
    g = Graph.bundled(name)
    step, prod, scales = g.info()
    oracle.graph_use(g)
    ostep, oprod, oscales, _, _ = oracle.graph_props(g.bounding_radius_hint)
    assert step == ostep and prod == oprod and scales == oscales


def test_bundled_graphs_are_the_reference_files():
    if not os.path.isdir(os.path.join(REF, "graphs")):
        pytest.skip("the reference tree is not present")
    for name in GRAPHS:
        with open(os.path.join(REF, "graphs", name + ".ron"), encoding="utf-8") as fh:
            assert Graph.from_ron(fh.read()) == Graph.bundled(name), name


def test_graph_round_trips_through_its_dict():
    for name in GRAPHS:
        g = Graph.bundled(name)
        assert Graph.from_dict(g.to_dict()) == g


def test_set_param_follows_the_bounds():
    g = Graph.bundled("mandelbulb")
    g.set_param("fractal.nodes[1].iterations", 10)
    g.set_param("fractal.nodes[1].bailout", 2.0)
    assert g.node(1).params["iterations"] == 10 and g.node(1).params["bailout"] == 2.0
    with pytest.raises(ValueError, match="outside"):
        g.set_param("fractal.nodes[1].bailout", 1.0)
    with pytest.raises(ValueError, match="no such parameter"):
        g.set_param("fractal.nodes[1].radius", 1.0)
    g = Graph.bundled("evidement")
    g.set_param("nodes[8].offset.y", 0.25)
    assert g.node(8).params["offset"] == (0.0, 0.25, 0.0)
    g.set_param("fractal.safety", 0.4)
    assert g.safety == 0.4


def _hiprtc():
    import ctypes
    path = "/opt/rocm/lib/libhiprtc.so"
    if not os.path.exists(path):
        pytest.skip("hiprtc is not installed")
    return ctypes.CDLL(path)


@pytest.mark.parametrize("name", GRAPHS)
def test_specialised_source_compiles(name):
    """The generated device source of every bundled topology compiles (its probe kernel, with the
    options the library passes to hiprtc at run time); the other kernels share its header code."""
    import ctypes
    import glob
    rtc = _hiprtc()
    src = Graph.bundled(name).source()
    assert "graph_jit" in src and "GOP_" in src
    csrc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "syfrack_amd", "csrc")
    clang = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/include"))[-1]
    opts = [b"--offload-arch=gfx950", b"-O1", b"-std=c++17", b"-ffp-contract=off", b"-w",
            ("-I" + csrc).encode(), b"-I/opt/rocm/include", ("-I" + clang).encode()]
    prog = ctypes.c_void_p()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b"t.hip", 0, None, None) == 0
    rtc.hiprtcAddNameExpression(prog, b"syf::k_de_probe<7u>")
    arr = (ctypes.c_char_p * len(opts))(*opts)
    rc = rtc.hiprtcCompileProgram(prog, len(opts), arr)
    n = ctypes.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
    log = ctypes.create_string_buffer(n.value + 1)
    rtc.hiprtcGetProgramLog(prog, log)
    rtc.hiprtcDestroyProgram(ctypes.byref(prog))
    assert rc == 0, log.value.decode(errors="replace")[:2000]


def test_graphs_of_one_topology_share_their_source():
    """The specialised source depends on the structure only: parameters are read at run time."""
    a = Graph.bundled("mandelbox")
    b = a.copy()
    b.set_param("fractal.nodes[1].iterations", 9)
    b.set_param("fractal.nodes[5].factor", -1.5)
    b.set_param("fractal.nodes[4].min_radius", 0.25)
    assert a.source() == b.source()
    c = a.copy()
    c.node(3).params["slot"] = 2  # a trap slot is structure
    assert a.source() != c.source()


def test_classic_topologies_run_on_the_classic_kernels():
    from syfrack_amd import abi
    fam = {"mandelbulb": abi.FAMILY_MANDELBULB, "mandelbox": abi.FAMILY_MANDELBOX, "menger": abi.FAMILY_MENGER,
           "sphere": abi.FAMILY_SPHERE}
    for name in GRAPHS:
        assert Graph.bundled(name).kernel_family() == fam.get(name, abi.FAMILY_GRAPH), name
    g = Graph.bundled("mandelbulb")
    g.set_param("fractal.nodes[1].iterations", 7)
    g.set_param("fractal.nodes[3].power", 5.5)
    assert g.kernel_family() == abi.FAMILY_MANDELBULB  # parameters do not matter
    for change in [lambda g: setattr(g.node(2), "enabled", False),            # a disabled box fold
                   lambda g: g.node(7).params.__setitem__("inset", 0.5),      # an inset surface
                   lambda g: g.node(3).params.__setitem__("slot", 1),         # another trap slot
                   lambda g: g.node(3).params.__setitem__("mode", "Last")]:   # another trap mode
        g = Graph.bundled("mandelbox")
        change(g)
        assert g.kernel_family() == abi.FAMILY_GRAPH
