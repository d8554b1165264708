"""Every workgroup barrier of every kernel sits in wave-uniform control flow (CPU only).

s_barrier counts waves, not lanes.  If the compiler takes a loop's exit condition for
divergent (say, a block index read back from LDS), it may structure the loop with exec
masks: a wave then keeps running the loop for some of its lanes and reaches the barrier on
their behalf while its other lanes still have work to do before that barrier -- for
example lane 0 drawing the next work ticket that the rest of the workgroup is about to
read.  Such a kernel reuses stale values or hangs.  Persistent kernels here keep the
values that steer their loops uniform (readfirstlane of the LDS copy, or a form the
compiler proves uniform); this test checks the result in the gfx950 assembly of every
kernel translation unit:

  for every loop (the compiler's own loop annotations) that contains an s_barrier, no
  branch inside the loop that tests exec (s_cbranch_execz / execnz) may leave the loop or
  jump back to its header -- i.e. every loop around a barrier is entered, repeated and left
  by all lanes of a wave together.
"""
import glob
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "opt-fft_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

LABEL = re.compile(r"^(\.LBB\d+_\d+):(.*)$")
FALL = re.compile(r"^; %bb\.\d+:(.*)$")
BRANCH = re.compile(r"\b(s_cbranch_\w+|s_branch)\s+\.L(BB\d+_\d+)")


def _blocks(body):
    """Split one function's lines into blocks: (name, annotation text, instruction lines)."""
    blocks, cur = [], None
    for ln in body:
        m, f = LABEL.match(ln), FALL.match(ln)
        if m or f:
            # ".LBB0_4" is block "BB0_4", the name the loop annotations use
            cur = [m.group(1)[2:] if m else None, (m or f).group(m and 2 or 1), []]
            blocks.append(cur)
        elif cur is not None:
            s = ln.strip()
            if s.startswith(";") and not cur[2]:
                cur[1] += " " + s
            else:
                cur[2].append(ln)
    return blocks


def _loop_tree(blocks):
    """innermost loop header of each block, and the parent of each loop header."""
    inner, parent = [], {}
    for name, ann, _ in blocks:
        h = None
        if "Loop Header: Depth=" in ann:
            h = name
            depth = int(re.search(r"Loop Header: Depth=(\d+)", ann).group(1))
            par = dict((int(d), p) for p, d in re.findall(r"Parent Loop (BB\d+_\d+) Depth=(\d+)", ann))
            parent[name] = par.get(depth - 1)
        else:
            m = re.search(r"Header=(BB\d+_\d+) Depth=\d+", ann)
            h = m.group(1) if m else None
        inner.append(h)
    return inner, parent


def _in_loop(h, loop, parent):
    while h is not None:
        if h == loop:
            return True
        h = parent.get(h)
    return False


def divergent_barrier_loops(asm_text):
    """[(kernel, loop header, offending branch)] for one assembly file."""
    lines = asm_text.split("\n")
    starts = [i for i, ln in enumerate(lines) if re.match(r"^_Z\S+:", ln)]
    bad = []
    for st in starts:
        end = st + 1
        while end < len(lines) and not lines[end].startswith(".Lfunc_end"):
            end += 1
        kname = lines[st][:-1]
        blocks = _blocks(lines[st + 1:end])
        inner, parent = _loop_tree(blocks)
        where = {b[0]: i for i, b in enumerate(blocks) if b[0]}
        loops = set(h for h in inner if h)
        for loop in loops:
            members = [i for i, h in enumerate(inner) if _in_loop(h, loop, parent)]
            if not any("s_barrier" in ln for i in members for ln in blocks[i][2]):
                continue
            for i in members:
                for ln in blocks[i][2]:
                    m = BRANCH.search(ln)
                    if not m or "exec" not in m.group(1):
                        continue
                    tgt = where.get(m.group(2))
                    leaves = tgt is None or not _in_loop(inner[tgt], loop, parent)
                    if leaves or m.group(2) == loop:
                        bad.append((kname, loop, ln.strip()))
    return bad


@pytest.fixture(scope="module")
def assembly():
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    assert len(srcs) >= 6
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        procs = []
        for s in srcs:
            dst = os.path.join(tmp, os.path.basename(s) + ".s")
            cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                   "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only",
                   "-S", "-I" + CSRC, "-I" + os.path.join(ROOT, "include"), "-o", dst, s]
            procs.append((s, dst, subprocess.Popen(cmd, stdout=subprocess.DEVNULL,
                                                   stderr=subprocess.PIPE)))
        for s, dst, p in procs:
            _, err = p.communicate()
            assert p.returncode == 0, f"{s}: {err.decode()[-2000:]}"
            out[os.path.basename(s)] = open(dst).read()
    return out


def test_barriers_only_in_uniform_loops(assembly):
    nkern = 0
    bad = []
    for name, text in assembly.items():
        nkern += len(re.findall(r"^_Z\S+:", text, re.M))
        bad += [(name,) + b for b in divergent_barrier_loops(text)]
    assert nkern >= 100
    assert not bad, "barrier inside a loop left or repeated under exec masks:\n" + "\n".join(
        f"{f}: {k[:90]} loop {h}: {br}" for f, k, h, br in bad[:10])


# A loop in the shape the compiler produced for a persistent kernel whose ticket it took for
# divergent: lanes that still have to draw the ticket leave the inner loop under an exec
# mask (s_cbranch_execz to the outer loop), the others go back to the barrier at its top.
_BAD = """\
_Zk_example:
.LBB0_4:                                ; =>This Loop Header: Depth=1
                                        ;     Child Loop BB0_14 Depth 2
\tglobal_atomic_add v3, v147, v3, s[14:15] sc0
\ts_branch .LBB0_14
.LBB0_13:                               ;   in Loop: Header=BB0_14 Depth=2
\ts_andn2_b64 exec, exec, s[58:59]
\ts_cbranch_execz .LBB0_42
.LBB0_14:                               ;   Parent Loop BB0_4 Depth=1
                                        ; =>  This Inner Loop Header: Depth=2
\ts_barrier
\tds_read_b64 v[2:3], v147 offset:63936
\ts_branch .LBB0_13
.LBB0_42:                               ;   in Loop: Header=BB0_4 Depth=1
\ts_cbranch_vccz .LBB0_4
\ts_endpgm
.Lfunc_end0:
"""

# The same loop with a uniform (scalar) exit: accepted.
_GOOD = _BAD.replace("s_cbranch_execz .LBB0_42", "s_cbranch_vccnz .LBB0_42")


def test_checker_flags_the_divergent_shape():
    assert divergent_barrier_loops(_BAD), "checker missed a barrier loop left under exec"
    assert not divergent_barrier_loops(_GOOD)
