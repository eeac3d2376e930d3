"""Restart files (CPU): the TorchScript payload format, the bundle index and
the restart output schedule of MeshBlock.

The payload format belongs to kintera (an unvendored dependency of the
reference).  It is pinned from both sides the reference's own Python code
fixes: python/api/pd_inspect.py:21-60 reads a payload with torch.jit.load and
named_buffers, and python/write_init.py:12-22 writes one by scripting an
nn.Module whose buffers are the tensors.  The bundle layout is pinned by
bytes assembled here from src/output/combine_restart.cpp:47-74 and the
reader's rules in src/input/read_restart_file.cpp:36-235."""
import os
import warnings

import pytest
import torch

from snapy_linfel_amd import host

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class TensorModule(torch.nn.Module):
    """The payload writer of python/write_init.py:12-22."""

    def __init__(self, tensors):
        super().__init__()
        for name, tensor in tensors.items():
            self.register_buffer(name, tensor)


def _script_save(tensors, filename):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        torch.jit.script(TensorModule(tensors)).save(filename)


def _sample(seed):
    g = torch.Generator().manual_seed(seed)
    return {
        "hydro_u": torch.randn(5, 1, 8, 14, generator=g, dtype=torch.float64),
        "hydro_w": torch.randn(5, 1, 8, 14, generator=g, dtype=torch.float32),
        "last_time": torch.tensor([123.25], dtype=torch.float64),
        "last_cycle": torch.tensor([seed], dtype=torch.int64),
    }


def _same(a, b):
    assert sorted(a) == sorted(b)
    for k in a:
        assert a[k].dtype == b[k].dtype and torch.equal(a[k], b[k]), k


def test_payload_reads_as_torchscript_buffers(tmp_path):
    v = _sample(3)
    f = str(tmp_path / "x.block0.00000.part")
    host.save_tensors(v, f)
    _same(dict(torch.jit.load(f).named_buffers()), v)
    _same(host.load_tensors(f), v)
    _same(host.load_restart(f), v)


def test_reads_payloads_written_by_scripting(tmp_path):
    v = _sample(4)
    f = str(tmp_path / "init.pt")
    _script_save(v, f)
    _same(host.load_tensors(f), v)
    _same(host.load_restart(f, 0), v)


def _bundle_bytes(entries):
    """A bundle assembled from the format rules: magic, count, name<TAB>size
    lines, an empty line, then the payloads in index order."""
    head = "SNAPY_RESTART_BUNDLE_V1\n%d\n" % len(entries)
    head += "".join("%s\t%d\n" % (n, len(b)) for n, b in entries) + "\n"
    return head.encode() + b"".join(b for _, b in entries)


def _payload(tmp_path, name, v):
    f = str(tmp_path / name)
    host.save_tensors(v, f)
    return open(f, "rb").read()


def test_bundle_writer_layout(tmp_path):
    names = ["case.block0.00002.part", "case.block1.00002.part"]
    files = [str(tmp_path / n) for n in names]
    vs = [_sample(10), _sample(11)]
    for f, v in zip(files, vs):
        host.save_tensors(v, f)
    out = str(tmp_path / "case.00002.restart")
    assert host.make_restart_bundle(out, files) == 0
    expect = _bundle_bytes([(n, open(f, "rb").read()) for n, f in zip(names, files)])
    assert open(out, "rb").read() == expect
    _same(host.load_restart(out, 0), vs[0])
    _same(host.load_restart(out, 1), vs[1])
    assert host.make_restart_bundle(str(tmp_path / "y.restart"), [str(tmp_path / "missing")]) == -1


def test_bundle_reader_picks_the_block(tmp_path, capfd):
    vs = {r: _sample(20 + r) for r in (0, 1, 2)}
    entries = [("notes.txt", b"not a payload\n")]
    for r in (2, 0, 1):  # index order is not rank order
        entries.append(("run.block%d.00007.part" % r, _payload(tmp_path, "p%d" % r, vs[r])))
    f = tmp_path / "run.00007.restart"
    f.write_bytes(_bundle_bytes(entries))
    for r in (0, 1, 2):
        _same(host.load_restart(str(f), r), vs[r])
    # no payload for the block: empty result and a message, as the reference
    assert host.load_restart(str(f), 5) == {}
    assert "no matching block" in capfd.readouterr().err


def test_bundle_reader_errors(tmp_path, capfd):
    good = _payload(tmp_path, "p0", _sample(1))
    # index line without a tab
    f = tmp_path / "a.restart"
    f.write_bytes(b"SNAPY_RESTART_BUNDLE_V1\n1\nrun.block0.00000.part 12\n\n" + good)
    assert host.load_restart(str(f), 0) == {}
    assert "missing tab" in capfd.readouterr().err
    # payload shorter than its index entry
    f = tmp_path / "b.restart"
    f.write_bytes(_bundle_bytes([("run.block0.00000.part", good)])[:-10])
    assert host.load_restart(str(f), 0) == {}
    assert "truncated" in capfd.readouterr().err
    # a payload name without the block field
    f = tmp_path / "c.restart"
    f.write_bytes(_bundle_bytes([("run.part", good)]))
    with pytest.raises(ValueError):
        host.load_restart(str(f), 0)
    # neither a bundle nor a payload
    f = tmp_path / "d.restart"
    f.write_bytes(b"plain text\n")
    with pytest.raises(RuntimeError):
        host.load_restart(str(f), 0)


def _meshblock(tmp_path, outputs):
    op = host.MeshBlockOptions.from_yaml(os.path.join(G, "sod.yaml"))
    outs = []
    for i, (kind, dt) in enumerate(outputs):
        o = host.OutputOptions()
        o.file_type, o.dt, o.fid = kind, dt, i
        outs.append(o)
    op.outputs = outs
    op.output_dir = str(tmp_path / "out")
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return host.MeshBlock(op)


def test_options_basename_and_outputs():
    op = host.MeshBlockOptions.from_yaml(os.path.join(G, "straka.yaml"))
    assert op.basename == "straka" and op.output_dir == "."
    assert [(o.file_type, o.dt, o.fid) for o in op.outputs] == [("restart", 300.0, 0),
                                                                 ("netcdf", 300.0, 1)]
    assert op.outputs[1].variables == ["prim", "uov"]


def test_restart_output_schedule(tmp_path):
    """make_outputs (src/mesh/meshblock.cpp:813-830) with the file contents of
    src/output/restart.cpp:25-86: the schedules a file records are the ones
    in force when it was written (its own output not yet advanced)."""
    mb = _meshblock(tmp_path, [("netcdf", 5.0), ("restart", 10.0)])
    out = tmp_path / "out"
    u = torch.arange(5 * 8, dtype=torch.float64).reshape(5, 1, 1, 8)
    times = [0.0, 4.0, 9.5, 10.0, 17.0, 21.0]
    for c, t in enumerate(times):
        mb.cycle = c
        mb.make_outputs({"hydro_u": u + c}, t)
    assert sorted(os.listdir(out)) == ["sod.00000.restart", "sod.00001.restart",
                                       "sod.00002.restart"]
    # written at t = 0 (cycle 0), 10 (cycle 3) and 21 (cycle 5)
    expect = {0: (0.0, 0, [1, 0], [5.0, 0.0]), 1: (10.0, 3, [3, 1], [15.0, 10.0]),
              2: (21.0, 5, [5, 2], [25.0, 20.0])}
    for n, (t, c, numbers, nexts) in expect.items():
        d = host.load_restart(str(out / ("sod.%05d.restart" % n)))
        assert d["last_time"].tolist() == [t] and d["last_cycle"].tolist() == [c]
        assert d["file_number"].tolist() == numbers and d["next_time"].tolist() == nexts
        assert torch.equal(d["hydro_u"], u + c)
    assert [(o.file_number, o.next_time) for o in mb.get_outputs()] == [(5, 25.0), (3, 30.0)]
    mb.finalize({"hydro_u": u}, 22.0)
    assert "sod.final.restart" in os.listdir(out)
    assert [(o.file_number, o.next_time) for o in mb.get_outputs()] == [(5, 25.0), (3, 30.0)]


def test_restart_output_combines_block_parts(tmp_path):
    """Parts of other blocks beside this block's are bundled and removed
    (src/output/combine_restart.cpp:77-170)."""
    mb = _meshblock(tmp_path, [("restart", 1.0)])
    out = tmp_path / "out"
    out.mkdir()
    other = _sample(9)
    host.save_tensors(other, str(out / "sod.block1.00000.part"))
    u = torch.ones(5, 1, 1, 8, dtype=torch.float64)
    mb.make_outputs({"hydro_u": u}, 0.0)
    assert os.listdir(out) == ["sod.00000.restart"]
    _same(host.load_restart(str(out / "sod.00000.restart"), 1), other)
    assert torch.equal(host.load_restart(str(out / "sod.00000.restart"), 0)["hydro_u"], u)


def test_unknown_output_type_and_missing_variables(tmp_path):
    with pytest.raises(RuntimeError, match="is not implemented"):
        _meshblock(tmp_path, [("hdf5", 1.0)])
    mb = _meshblock(tmp_path, [("restart", 1.0)])
    f = str(tmp_path / "partial.part")
    host.save_tensors({"hydro_u": torch.zeros(2), "last_time": torch.zeros(1)}, f)
    with pytest.raises(RuntimeError, match="missing required variable: last_cycle"):
        mb.initialize_from_restart(f)
