"""Message hooks (Cubiquity's set_debug_handler / set_warning_handler / setProgressHandler,
src/library/base.h:102-107): errors reach the warning handler with the text cqx_last_error
returns; volume generation reports progress from the first step to the last and debug notes."""
import pytest


@pytest.fixture
def hooks(cq):
    got = {"debug": [], "warning": [], "progress": []}
    cq.set_debug_handler(lambda m: got["debug"].append(m))
    cq.set_warning_handler(lambda m: got["warning"].append(m))
    cq.set_progress_handler(lambda *a: got["progress"].append(a))
    yield got
    cq.set_debug_handler(None)
    cq.set_warning_handler(None)
    cq.set_progress_handler(None)


def test_errors_reach_the_warning_handler(cq, hooks, tmp_path):
    with pytest.raises(cq.CqxError):
        cq.HostVolume.load_dag(str(tmp_path / "missing.dag"))
    assert hooks["warning"] and hooks["warning"][-1] == cq.cabi.lib().cqx_last_error().decode()
    assert "missing.dag" in hooks["warning"][-1]


def test_generation_reports_progress(cq, hooks):
    hv = cq.HostVolume.generate(cq.gen_default_params(9, 2))
    prog = [p for p in hooks["progress"] if p[0] == "Building volume"]
    assert len(prog) >= 2
    first, last = prog[0], prog[-1]
    assert first[1] == first[2] == 0 and last[2] == last[3] > 0
    assert all(a[2] <= b[2] for a, b in zip(prog, prog[1:]))
    assert any(m.startswith("built ") for m in hooks["debug"])
    assert any(m == f"renumbering {hv.node_count} nodes" for m in hooks["debug"])


def test_handlers_can_be_cleared(cq, tmp_path):
    got = []
    cq.set_warning_handler(got.append)
    cq.set_warning_handler(None)
    with pytest.raises(cq.CqxError):
        cq.HostVolume.load_dag(str(tmp_path / "missing.dag"))
    assert got == []
