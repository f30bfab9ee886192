"""Writes les_amd_capi.f90, the iso_c_binding interface of every function and
struct that include/les_amd.h declares, so Fortran callers of the reference
(whose code base is Fortran) bind the whole C ABI without hand-written
prototypes.  Run after a change to the header:

    python les_amd/fortran/gen_capi.py            # rewrite the module
    python les_amd/fortran/gen_capi.py --check    # exit 1 if it is stale

Mapping: numbers by value (uint64_t as integer(c_int64_t)); every data
pointer (device or host) as type(c_ptr), value; arrays of pointers as
type(c_ptr) :: x(*); the les_* structs as bind(C) derived types passed by
reference; les_ctx / les_ensemble handles as type(c_ptr); strings as
character(kind=c_char) :: x(*) (NUL-terminated by the caller).
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "..", "include", "les_amd.h")
OUT = os.path.join(HERE, "les_amd_capi.f90")

SCALARS = {"int": "integer(c_int)", "double": "real(c_double)", "float": "real(c_float)",
           "int64_t": "integer(c_int64_t)", "uint64_t": "integer(c_int64_t)", "size_t": "integer(c_size_t)"}
HANDLES = ("les_ctx", "les_ensemble")


def strip_comments(txt):
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return re.sub(r"//[^\n]*", "", txt)


def parse(txt):
    txt = strip_comments(txt)
    structs = []
    for tag, body, name in re.findall(r"typedef struct (\w+)\s*\{(.*?)\}\s*(\w+)\s*;", txt, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"((?:const )?\w+)\s*(.*)", decl)
            base, rest = m.group(1), m.group(2)
            for d in rest.split(","):
                d = d.strip()
                fields.append((base.replace("const ", ""), d.count("*") > 0, d.replace("*", "").strip()))
        structs.append((name, fields))
    funcs = []
    for ret, name, args in re.findall(r"([A-Za-z_][\w \*]*?)\b(les_\w+)\s*\(([^;{)]*)\)\s*;", txt, flags=re.S):
        ret = " ".join(ret.split())
        params = []
        a = " ".join(args.split())
        if a and a != "void":
            for p in a.split(","):
                p = p.strip()
                m = re.match(r"(.*?)(\w+)\s*(\[\d*\])?$", p)
                params.append((" ".join(m.group(1).split()), m.group(2), m.group(3) is not None))
        funcs.append((ret, name, params))
    return structs, funcs


def ftype(ctype, is_array, struct_names):
    t = ctype.replace("const ", "").replace(" ", "")
    if is_array:
        return ("type(c_ptr) :: {}(*)" if "*" in t else SCALARS[t.rstrip("*")] + " :: {}(*)"), False
    if t in SCALARS:
        return SCALARS[t] + ", value :: {}", False
    if t in ("char*",):
        return "character(kind=c_char) :: {}(*)", False
    base = t.rstrip("*")
    stars = len(t) - len(base)
    if base in HANDLES:
        return ("type(c_ptr), value :: {}" if stars == 1 else
                ("type(c_ptr) :: {}(*)" if "*const*" in ctype.replace(" ", "") else "type(c_ptr) :: {}")), False
    if base in struct_names and stars == 1:
        return "type({}) :: ".format(base) + "{}", True
    if stars == 2:
        return ("type(c_ptr) :: {}(*)" if "*const*" in ctype.replace(" ", "") else "type(c_ptr) :: {}"), False
    return "type(c_ptr), value :: {}", False


def wrap(line, indent):
    """free-form lines of at most 100 characters, continued with &"""
    out, cur = [], line
    while len(cur) > 100:
        cut = cur.rfind(",", 0, 98) + 1
        out.append(cur[:cut] + " &")
        cur = " " * (indent + 4) + cur[cut:].lstrip()
    out.append(cur)
    return "\n".join(out)


def generate():
    structs, funcs = parse(open(HEADER).read())
    names = {s[0] for s in structs}
    L = ["! les_amd_capi.f90 -- generated by gen_capi.py from include/les_amd.h; do not edit.",
         "! iso_c_binding interfaces of the whole les_amd C ABI (one per function).",
         "module les_amd_capi", "  use iso_c_binding", "  implicit none", ""]
    for name, fields in structs:
        L.append(f"  type, bind(C) :: {name}")
        for base, ptr, fname in fields:
            L.append("    " + ("type(c_ptr)" if ptr else SCALARS[base]) + f" :: {fname}")
        L.append(f"  end type {name}")
        L.append("")
    L.append("  interface")
    for ret, name, params in funcs:
        seen, dummies, decls = set(), [], []
        for ctype, pname, arr in params:
            n = pname
            while n.lower() in seen or n.lower() == name.lower():
                n += "_"
            seen.add(n.lower())
            dummies.append(n)
            fmt, intent_in = ftype(ctype, arr, names)
            d = "      " + fmt.format(n)
            if intent_in and ctype.startswith("const "):
                d = d.replace(" :: ", ", intent(in) :: ")
            decls.append(d)
        arglist = ", ".join(dummies)
        r = ret.replace(" ", "")
        if r == "void":
            head, tail = f"    subroutine {name}({arglist}) bind(C, name='{name}')", "    end subroutine"
        elif r == "int":
            head, tail = f"    integer(c_int) function {name}({arglist}) bind(C, name='{name}')", "    end function"
        else:
            head, tail = f"    type(c_ptr) function {name}({arglist}) bind(C, name='{name}')", "    end function"
        L.append(wrap(head, 4))
        L.append("      import")
        L.append("      implicit none")
        L.extend(decls)
        L.append(tail)
    L.append("  end interface")
    L.append("end module les_amd_capi")
    return "\n".join(L) + "\n", [f[1] for f in funcs]


def main():
    text, _ = generate()
    if "--check" in sys.argv:
        cur = open(OUT).read() if os.path.exists(OUT) else ""
        if cur != text:
            print("les_amd_capi.f90 is stale: run python les_amd/fortran/gen_capi.py")
            sys.exit(1)
        return
    with open(OUT, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
