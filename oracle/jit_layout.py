"""Layout probe of the JIT element-table ABI (test infrastructure).

Generates a C program that prints sizeof of every struct of include/pyoomph_jit_table.h
and the offsetof of every one of their fields, and builds it against a header that
declares those structs: this library's header, or the reference's src/jitbridge.h
(src/jitbridge.h:88-497), whose structs carry the same typedef names. Equal output means
the two layouts are the same, field by field. The field list is read from this library's
header after the C preprocessor has expanded it.

  python oracle/jit_layout.py own       -> oracle/_build/jit_layout_own, runs it
  python oracle/jit_layout.py ref       -> oracle/_ref/jit_layout_ref (needs the
                                           reference tree), runs it
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OWN_HEADER = os.path.join(ROOT, "include", "pyoomph_jit_table.h")
REF_HEADER = "/root/reference/src/jitbridge.h"
STRUCTS = ["JITElementInfo_t", "JITHangInfoEntry_t", "JITHangInfo_t", "JITShapeInfo_t",
           "JITFuncSpec_RequiredShapes_FiniteElement_t", "JITFuncSpec_Callback_Entry_t",
           "JITFuncSpec_MultiRet_Entry_t", "JITFuncSpec_Table_FiniteElement_t"]


def struct_fields():
    """{typedef name: [field names]} from the preprocessed own header."""
    pre = subprocess.run(["gcc", "-E", "-P", "-x", "c", OWN_HEADER], check=True, capture_output=True,
                         text=True).stdout
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+\w+\s*\{(.*?)\}\s*(\w+)\s*;", pre, flags=re.S):
        body, name = m.group(1), m.group(2)
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            fp = re.search(r"\(\s*\*\s*(\w+)\s*\)", decl)
            if fp:
                fields.append(fp.group(1))
                continue
            for part in decl.split(","):
                part = re.sub(r"\[.*?\]", "", part).strip()
                fields.append(re.findall(r"(\w+)\s*$", part.replace("*", " "))[0])
        out[name] = fields
    return out


def probe_source(header):
    fields = struct_fields()
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "%s"' % header, "int main(void) {"]
    for s in STRUCTS:
        lines.append('    printf("%s sizeof %%zu\\n", sizeof(%s));' % (s, s))
        for f in fields[s]:
            lines.append('    printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (s, f, s, f))
    lines += ["    return 0;", "}"]
    return "\n".join(lines) + "\n"


def build(which):
    if which == "own":
        header, outdir, exe = OWN_HEADER, os.path.join(HERE, "_build"), "jit_layout_own"
    else:
        header, outdir, exe = REF_HEADER, os.path.join(HERE, "_ref"), "jit_layout_ref"
        if not os.path.exists(header):
            raise FileNotFoundError(header)
    os.makedirs(outdir, exist_ok=True)
    src = os.path.join(outdir, exe + ".c")
    with open(src, "w") as f:
        f.write(probe_source(header))
    path = os.path.join(outdir, exe)
    subprocess.run(["gcc", "-O0", "-std=gnu99", "-w", "-o", path, src, "-lm"], check=True)
    return path


def layout(which):
    """{"Struct.field": offset, "Struct sizeof": size} of the built probe."""
    exe = build(which)
    res = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {line.rsplit(" ", 1)[0]: int(line.rsplit(" ", 1)[1]) for line in res.splitlines() if line}


if __name__ == "__main__":
    for k, v in layout(sys.argv[1] if len(sys.argv) > 1 else "own").items():
        print(k, v)
