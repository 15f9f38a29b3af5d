"""Static check of the built library's gfx950 code: every DPP instruction
reads its source register at least two wait states after the last VALU write
of that register (the hardware does not interlock this case; the CDNA3/CDNA4
ISA's table of manually inserted wait states requires 2).

The team kernels issue their broadcast fmas from inline asm, where the
compiler's hazard recognizer does not look; some of them leave the wait
states to a preceding s_nop over the source registers instead of one per
instruction. This test disassembles the code objects of libcbk_amd.so (this
repository's own build) and checks the rule on every path into every DPP
instruction, across branches.
"""
import os
import re
import shutil
import struct
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cbk_amd", "libcbk_amd.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
NEED = 2

_REG = re.compile(r"^v(\d+)$|^v\[(\d+):(\d+)\]$")
_INSN = re.compile(r"^\s+(\S+)\s*(.*?)\s*//\s*([0-9A-Fa-f]+):")
_FUNC = re.compile(r"^([0-9a-f]+) <(.+)>:$")
_TARGET = re.compile(r"<(.+)\+0x([0-9a-f]+)>\s*$")


def _regs(tok):
    tok = tok.strip().lstrip("-").strip("|")
    m = _REG.match(tok)
    if not m:
        return set()
    if m.group(1) is not None:
        return {int(m.group(1))}
    return set(range(int(m.group(2)), int(m.group(3)) + 1))


def code_objects(lib):
    """the gfx950 code objects of every bundle in the .hip_fatbin section"""
    tmp = tempfile.mkdtemp()
    try:
        fat = os.path.join(tmp, "fat.bin")
        subprocess.run([f"{LLVM}/llvm-objcopy", "--dump-section", f".hip_fatbin={fat}", lib,
                        os.path.join(tmp, "stripped")], check=True, capture_output=True)
        data = open(fat, "rb").read()
    finally:
        shutil.rmtree(tmp)
    out = []
    for m in re.finditer(re.escape(MAGIC), data):
        s = m.start()
        p = s + len(MAGIC)
        (n,) = struct.unpack_from("<Q", data, p)
        p += 8
        for _ in range(n):
            off, size, tl = struct.unpack_from("<QQQ", data, p)
            p += 24
            triple = data[p:p + tl].decode()
            p += tl
            if "gfx950" in triple:
                out.append(data[s + off:s + off + size])
    return out


def disassemble(blob):
    with tempfile.NamedTemporaryFile(suffix=".o") as f:
        f.write(blob)
        f.flush()
        r = subprocess.run([f"{LLVM}/llvm-objdump", "-d", "--mcpu=gfx950", f.name],
                           capture_output=True, text=True, check=True)
    return r.stdout


def parse(text):
    """functions -> list of (addr, mnemonic, operands, branch target addr)"""
    funcs = {}
    cur = None
    start = {}
    for line in text.splitlines():
        fm = _FUNC.match(line)
        if fm:
            cur = fm.group(2)
            start[cur] = int(fm.group(1), 16)
            funcs[cur] = []
            continue
        im = _INSN.match(line)
        if im and cur is not None:
            mnem, ops, addr = im.group(1), im.group(2), int(im.group(3), 16)
            tgt = None
            if mnem.startswith("s_branch") or mnem.startswith("s_cbranch"):
                tm = _TARGET.search(line)
                if tm and tm.group(1) in start:
                    tgt = start[tm.group(1)] + int(tm.group(2), 16)
            funcs[cur].append((addr, mnem, ops, tgt))
    return funcs


def _valu_writes(mnem, ops):
    if not mnem.startswith("v_") or mnem.startswith(("v_cmp", "v_readlane", "v_readfirstlane")):
        return set()
    first = ops.split(",")[0] if ops else ""
    return _regs(first)


def _dpp_source(mnem, ops):
    if not mnem.endswith("_dpp"):
        return set()
    parts = [p.strip() for p in ops.split(",")]
    if len(parts) < 2:
        return set()
    return _regs(parts[1].split()[0])


def _waits(mnem, ops):
    if mnem == "s_nop":
        return int(ops.split()[0], 0) + 1
    return 1


def hazards(insns):
    index = {a: i for i, (a, _, _, _) in enumerate(insns)}
    preds = {}
    for i, (a, m, o, t) in enumerate(insns):
        if t is not None and t in index:
            preds.setdefault(index[t], []).append(i)
    bad = []

    def search(k, regs, have, seen):
        """walk back from instruction k (exclusive) while fewer than NEED wait
        states separate it from the DPP read"""
        cands = []
        if k > 0 and not insns[k - 1][1] in ("s_branch", "s_endpgm", "s_setpc_b64"):
            cands.append(k - 1)
        cands += preds.get(k, [])
        for j in cands:
            if (j, have) in seen:
                continue
            seen.add((j, have))
            _, m, o, _ = insns[j]
            if _valu_writes(m, o) & regs:
                return j
            h = have + _waits(m, o)
            if h < NEED:
                r = search(j, regs, h, seen)
                if r is not None:
                    return r
        return None

    for k, (a, m, o, _) in enumerate(insns):
        src = _dpp_source(m, o)
        if not src:
            continue
        j = search(k, src, 0, set())
        if j is not None:
            bad.append((insns[j], insns[k]))
    return bad


@pytest.mark.skipif(not os.path.exists(LIB), reason="libcbk_amd.so not built")
@pytest.mark.skipif(not os.path.exists(f"{LLVM}/llvm-objdump"), reason="no ROCm llvm-objdump")
def test_dpp_source_wait_states():
    blobs = code_objects(LIB)
    assert blobs, "no gfx950 code objects in the library"
    found = []
    ndpp = 0
    for blob in blobs:
        for name, insns in parse(disassemble(blob)).items():
            ndpp += sum(1 for _, m, _, _ in insns if m.endswith("_dpp"))
            for w, r in hazards(insns):
                found.append(f"{name}: {w[1]} {w[2]} @{w[0]:#x} -> {r[1]} {r[2]} @{r[0]:#x}")
    assert ndpp > 0
    assert not found, f"{len(found)} DPP reads too close to a VALU write:\n" + "\n".join(found[:20])


def test_checker_finds_a_planted_hazard():
    """the checker itself: a DPP read right after the write of its source, with
    and without the wait states, and across a branch"""
    ok = [(0, "v_fma_f64", "v[2:3], v[4:5], v[6:7], v[2:3]", None),
          (8, "s_nop", "1", None),
          (12, "v_fmac_f64_dpp", "v[0:1], -v[2:3], v[8:9] row_newbcast:1", None)]
    assert hazards(ok) == []
    bad = [(0, "v_fma_f64", "v[2:3], v[4:5], v[6:7], v[2:3]", None),
           (8, "v_add_f64", "v[10:11], v[4:5], v[6:7]", None),
           (16, "v_fmac_f64_dpp", "v[0:1], -v[3:4], v[8:9] row_newbcast:1", None)]
    assert len(hazards(bad)) == 1
    branch = [(0, "v_mov_b32_e32", "v5, 0", None),
              (4, "s_branch", "1", 12),
              (8, "s_nop", "7", None),
              (12, "v_mov_b32_dpp", "v6, v5 row_newbcast:0", None)]
    assert len(hazards(branch)) == 1
