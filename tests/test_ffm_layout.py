"""The Java FFM struct layouts written in INTEGRATION.md describe the compiled
structs of include/amoeba_nb.h: same fields, in the same order, at the same
byte offsets, with the same total size.

FFM's MemoryLayout.structLayout does not insert padding; it rejects a member
whose offset is not a multiple of its alignment.  The layouts are parsed from
the document's text, laid out by those rules here, and compared with the
ctypes mirrors and with nb_struct_sizes() of the built library.
"""
import ctypes
import os
import re

import pytest

from forcefieldx_amd import ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# FFM value layouts on x86-64 / aarch64 Linux: (size, alignment)
_VALUE = {"JAVA_DOUBLE": (8, 8), "JAVA_LONG": (8, 8), "JAVA_INT": (4, 4), "ADDRESS": (8, 8)}


def _split_top(s):
    """Split a comma-separated argument list at nesting depth zero."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        out.append("".join(cur).strip())
    return out


def java_layouts():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    layouts = {}
    for m in re.finditer(r"StructLayout\s+(\w+)\s*=\s*MemoryLayout\.structLayout\(", text):
        i, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        layouts[m.group(1)] = _split_top(text[m.end():i - 1])
    return layouts


def lay_out(members):
    """(name, offset, size) of each named member, and the struct size, by FFM's rules."""
    off, align_max, fields = 0, 1, []
    for mem in members:
        mem = " ".join(mem.split())
        pad = re.fullmatch(r"MemoryLayout\.paddingLayout\((\d+)\)", mem)
        if pad:
            off += int(pad.group(1))
            continue
        seq = re.fullmatch(r'MemoryLayout\.sequenceLayout\((\d+),\s*(\w+)\)\.withName\("(\w+)"\)', mem)
        val = re.fullmatch(r'(\w+)\.withName\("(\w+)"\)', mem)
        if seq:
            count, kind, name = int(seq.group(1)), seq.group(2), seq.group(3)
        elif val:
            count, kind, name = 1, val.group(1), val.group(2)
        else:
            raise AssertionError(f"unparsed layout member: {mem}")
        size, align = _VALUE[kind]
        # FFM throws IllegalArgumentException for a misaligned member
        assert off % align == 0, f"{name} at offset {off} is not {align}-aligned"
        fields.append((name, off, size * count))
        off += size * count
        align_max = max(align_max, align)
    assert off % align_max == 0, "struct size is not a multiple of its alignment"
    return fields, off


def ctypes_fields(struct):
    return [(name.rstrip("_"), getattr(struct, name).offset, ctypes.sizeof(t)) for name, t in struct._fields_]


@pytest.mark.parametrize("java_name, struct, slot", [("PARAMS", ffi.NbParams, 0), ("TOPOLOGY", ffi.NbTopology, 1),
                                                    ("ENERGIES", ffi.NbEnergies, 2)])
def test_java_layout_matches_compiled_struct(java_name, struct, slot):
    layouts = java_layouts()
    assert java_name in layouts, f"INTEGRATION.md has no {java_name} layout"
    fields, size = lay_out(layouts[java_name])
    assert fields == ctypes_fields(struct)
    L = ffi.load()
    sizes = (ctypes.c_int32 * 3)()
    assert L.nb_struct_sizes(sizes) == ffi.NB_OK
    assert size == sizes[slot] == ctypes.sizeof(struct)


def test_layout_checker_rejects_a_wrong_layout():
    # the first binding declared 15 leading doubles: every later field moved by 16 B
    bad = ['MemoryLayout.sequenceLayout(15, JAVA_DOUBLE).withName("terms")', 'JAVA_INT.withName("pcg_iters")',
           "MemoryLayout.paddingLayout(4)", 'JAVA_DOUBLE.withName("pcg_eps")']
    fields, _ = lay_out(bad)
    assert dict((n, o) for n, o, _ in fields)["pcg_iters"] == 120
    assert ffi.NbEnergies.pcg_iters.offset == 104
    with pytest.raises(AssertionError):
        lay_out(['JAVA_INT.withName("a")', 'JAVA_DOUBLE.withName("b")'])
