"""Source converter: a C MPI program -> a rank program of this runtime.

Follows the reference's source_converter/converter.cpp:27-313 (a clang
libTooling rewriter), here on clang's JSON AST dump instead of libTooling:

  * every function defined or declared in the program becomes __device__;
    `main` becomes `__gpu_main(int argc, char** argv)` (converter.cpp:49-65);
  * every variable with static storage becomes a per-rank array
    `T var[GPU_MPI_MAX_RANKS]` and every use `__gpu_global(var)`, so ranks do
    not share program globals (converter.cpp:113-160; `__gpu_global` indexes
    by rank here, kernel.hpp); read-only globals (top-level const) stay one
    __device__ copy;
  * C's implicit pointer conversions the C++ dialect refuses (void* -> T*,
    int -> enum) get explicit casts (converter.cpp:103-107);
  * identifiers that are C++ keywords are renamed `__decl_<name>`
    (converter.cpp:166-171 did this for `class`);
  * the program gets gpumpi/converted.h first (the C headers, then the MPI,
    stdio and libc mappings of the device) and gpumpi/gpu_main.h last (the
    rank kernel and a host main() with the `---gpumpi` command line).

The program is parsed as C against a C view of include/gpumpi/mpi.h that this
module derives from that header, so the MPI surface has one definition.

    python dphpc-gpu_mpi_amd/converter.py prog.c -o prog.cpp [-I dir ...]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CLANG = os.environ.get("GPUMPI_CLANG", "/opt/rocm/llvm/bin/clang")
MAX_RANKS = int(os.environ.get("GPU_MPI_MAX_RANKS", "1024"))

CXX_KEYWORDS = {"class", "new", "delete", "this", "template", "typename", "namespace", "operator", "private",
                "protected", "public", "virtual", "friend", "mutable", "explicit", "export", "try", "catch",
                "throw", "using", "nullptr", "constexpr", "decltype", "noexcept", "static_assert", "thread_local",
                "alignof", "char16_t", "char32_t", "concept", "requires", "co_await", "co_return", "co_yield"}


def c_view_of_mpi_h(text):
    """include/gpumpi/mpi.h as a C header (declarations only)."""
    out = []
    in_status = in_enum = False
    for line in text.splitlines():
        if "#include <hip/hip_runtime.h>" in line or "gpumpi_rt/mpi_impl.hpp" in line:
            continue
        if "operator==" in line or "operator!=" in line:
            continue
        if line.startswith("struct MPI_Status {"):
            in_status = True
        if in_status:
            line = re.sub(r"\s*=\s*[^;]+;", ";", line)
            if line.startswith("};"):
                in_status = False
                out.append(line)
                out.append("typedef struct MPI_Status MPI_Status;")
                continue
        if line.startswith("enum MPI_Datatype {"):
            in_enum = True
        line = line.replace("__device__ ", "")
        line = line.replace("nullptr", "((void*)0)")
        line = re.sub(r"\(MPI_Info\{0\}\)", "((MPI_Info){0})", line)
        line = re.sub(r"\(MPI_File\{([^}]*)\}\)", r"((MPI_File){\1})", line)
        out.append(line)
        if in_enum and line.startswith("};"):
            in_enum = False
            out.append("typedef enum MPI_Datatype MPI_Datatype;")
    return "\n".join(out) + "\n"


def parse_ast(src, include_dirs, tmp):
    mpi_h = os.path.join(ROOT, "include", "gpumpi", "mpi.h")
    with open(mpi_h) as f:
        c_h = c_view_of_mpi_h(f.read())
    inc = os.path.join(tmp, "include")
    os.makedirs(inc, exist_ok=True)
    with open(os.path.join(inc, "mpi.h"), "w") as f:
        f.write(c_h)
    cmd = [CLANG, "-fsyntax-only", "-x", "c", "-std=gnu11", "-Wno-everything", "-I", inc]
    for d in include_dirs:
        cmd += ["-I", d]
    cmd += ["-Xclang", "-ast-dump=json", src]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"parsing {src} failed:\n{p.stderr}")
    return json.loads(p.stdout)


class Locator:
    """clang's JSON dump names a location's file only when it changes (in
    dump order); this replays that order to give every location its file."""

    def __init__(self):
        self.file = None
        self.spelled = None  # spelling location of the latest macro-expanded loc

    def see(self, loc):
        """(file, offset, tokLen) of a location outside macro expansions,
        else None; for a macro-expanded one, self.spelled is where its token
        is written (a macro argument, or a #define body)."""
        self.spelled = None
        if not isinstance(loc, dict):
            return None
        spelled = None
        for key in ("spellingLoc", "expansionLoc"):
            if key in loc:
                inner = self.see(loc[key])
                if key == "spellingLoc":
                    spelled = inner
        if "file" in loc:
            self.file = loc["file"]
        self.spelled = spelled
        if "offset" in loc and "spellingLoc" not in loc:
            return (self.file, loc["offset"], loc.get("tokLen", 0))
        return None


def collect(ast, main_file, view_dir=None):
    """Edits as (offset, order, kind, payload) on the main file."""
    main_real = os.path.realpath(main_file)
    lc = Locator()
    nodes = []

    spellings = {}  # id(node) -> (begin, end) spelling locations of a macro-expanded node

    def visit(n, parents):
        loc = lc.see(n.get("loc"))
        rng = n.get("range", {})
        b = lc.see(rng.get("begin"))
        bs = lc.spelled
        e = lc.see(rng.get("end"))
        es = lc.spelled
        if b is None and bs is not None and es is not None:
            spellings[id(n)] = (bs, es)
        nodes.append((n, parents, loc, b, e))
        for c in n.get("inner", []):
            visit(c, parents + [n])

    visit(ast, [])

    def in_main(l):
        return l is not None and l[0] is not None and os.path.realpath(l[0]) == main_real

    globals_ = {}  # id -> VarDecl node of wrapped (per-rank) variables
    edits = []
    for n, parents, loc, b, e in nodes:
        k = n.get("kind")
        if k == "VarDecl" and in_main(loc) and in_main(b):
            file_scope = parents and parents[-1].get("kind") == "TranslationUnitDecl"
            static_local = not file_scope and n.get("storageClass") == "static"
            if not (file_scope or static_local):
                continue
            qt = n.get("type", {}).get("qualType", "")
            top_const = (qt.startswith("const ") and "*" not in qt) or qt.replace(" ", "").endswith("*const")
            if file_scope:
                edits.append((b[1], 0, "insert", "__device__ "))
            if top_const:
                continue
            globals_[n["id"]] = n
            edits.append((loc[1] + loc[2], 1, "insert", f"[{MAX_RANKS}]"))
            if n.get("init") == "c" and n.get("inner"):
                init = n["inner"][-1]
                ib = init.get("range", {}).get("begin", {})
                ie = init.get("range", {}).get("end", {})
                if "offset" in ib and "offset" in ie:
                    edits.append((ib["offset"], 0, "init", (ie["offset"] + ie.get("tokLen", 0))))
    # references, casts, functions, keyword renames
    for n, parents, loc, b, e in nodes:
        k = n.get("kind")
        inside_function = any(p.get("kind") == "FunctionDecl" for p in parents)
        if k == "DeclRefExpr" and inside_function and b is None and id(n) in spellings:
            # written inside a macro argument or a #define body of this file
            bs, es = spellings[id(n)]
            ref = n.get("referencedDecl", {})
            if in_main(bs) and in_main(es) and bs[1] == es[1] and ref.get("id") in globals_:
                edits.append((bs[1], 2, "insert", "__gpu_global("))
                edits.append((es[1] + es[2], 0, "insert", ")"))
        elif k == "DeclRefExpr" and in_main(b) and inside_function:
            ref = n.get("referencedDecl", {})
            if ref.get("id") in globals_:
                edits.append((b[1], 2, "insert", "__gpu_global("))
                edits.append((e[1] + e[2], 0, "insert", ")"))
            elif ref.get("name") in CXX_KEYWORDS:
                edits.append((b[1], 1, "insert", "__decl_"))
        elif k == "ImplicitCastExpr" and in_main(b) and inside_function:
            ck = n.get("castKind")
            qt = n.get("type", {}).get("qualType", "")
            if ck == "BitCast" or (ck == "IntegralCast" and qt.startswith("enum ")):
                edits.append((b[1], 1, "insert", f"({qt})"))
        elif k == "FunctionDecl" and in_main(loc) and in_main(b) and not n.get("isImplicit"):
            edits.append((b[1], 0, "insert", "__device__ "))
            if n.get("name") == "main":
                edits.append((loc[1], 0, "main", loc[1] + loc[2]))
            else:
                kr = kr_parameters(n, loc, main_file)
                if kr:
                    edits.append(kr)
        if k in ("VarDecl", "ParmVarDecl", "FunctionDecl") and in_main(loc) and n.get("name") in CXX_KEYWORDS:
            edits.append((loc[1], 1, "insert", "__decl_"))
    # Prototypes in the program's own headers (not converted themselves) get
    # a device redeclaration after the main file's last #include: clang
    # overloads on the target, so device code calls the device function.
    decls = []
    for n, parents, loc, b, e in nodes:
        if (n.get("kind") != "FunctionDecl" or n.get("isImplicit") or not loc or not loc[0] or in_main(loc)
                or not b or not e or b[0] != loc[0] or e[0] != loc[0] or not user_header(loc[0], view_dir)
                or not parents or parents[-1].get("kind") != "TranslationUnitDecl"
                or any(c.get("kind") == "CompoundStmt" for c in n.get("inner", []))):
            continue
        with open(loc[0], "rb") as f:
            text = f.read()[b[1]:e[1] + e[2]].decode()
        d = "__device__ " + " ".join(text.split()) + ";"
        if d not in decls:
            decls.append(d)
    if decls:
        with open(main_file, "rb") as f:
            data = f.read()
        at = 0
        for m in re.finditer(rb"^[ \t]*#[ \t]*include[^\n]*\n", data, re.M):
            at = m.end()
        edits.append((at, -1, "insert", "\n".join(decls) + "\n"))
    return edits


SYSTEM_PREFIXES = ("/usr/", "/opt/", "/lib/", os.path.join(ROOT, "include") + os.sep)


def user_header(path, view_dir=None):
    """A header of the program's own: not the C library's, not this
    runtime's, not the generated C view of mpi.h (in `view_dir`)."""
    real = os.path.realpath(path)
    pfx = SYSTEM_PREFIXES + ((os.path.realpath(view_dir) + os.sep,) if view_dir else ())
    return not any(real.startswith(p) for p in pfx)


def kr_parameters(fn, loc, main_file):
    """An old-style (K&R) definition -- `double f(X, A) double *X; double *A; {`
    -- as a prototype: the edit replacing everything from its parameter
    list to the body with `(double *X, double *A) ` (the reference's
    converter fixed these too)."""
    body = next((c for c in fn.get("inner", []) if c.get("kind") == "CompoundStmt"), None)
    if body is None:
        return None
    bb = body.get("range", {}).get("begin", {})
    if "offset" not in bb or "spellingLoc" in bb:
        return None
    with open(main_file, "rb") as f:
        data = f.read()
    start, stop = loc[1] + loc[2], bb["offset"]
    seg = data[start:stop].decode(errors="replace")
    i = seg.find("(")
    if i < 0:
        return None
    depth, j = 0, i
    while j < len(seg):
        depth += seg[j] == "("
        depth -= seg[j] == ")"
        if depth == 0:
            break
        j += 1
    rest = re.sub(r"/\*.*?\*/|//[^\n]*", "", seg[j + 1:], flags=re.S).strip()
    if not rest:
        return None  # a prototype already
    params = []
    for p in fn.get("inner", []):
        if p.get("kind") != "ParmVarDecl":
            continue
        t, name = p.get("type", {}).get("qualType", "int"), p.get("name", "")
        if "(" in t:
            return None  # function-pointer parameters stay as written
        params.append(f"{t} {name}")  # array parameters come adjusted to pointers
    return (start + i, 0, "replace", (stop, "(" + ", ".join(params) + ") "))


def apply(text, edits):
    data = text.encode()
    out = []
    pos = 0
    for off, order, kind, payload in sorted(set(edits), key=lambda x: (x[0], x[1], str(x[3]))):
        if off < pos:
            continue  # inside a span already replaced
        out.append(data[pos:off])
        pos = off
        if kind == "insert":
            out.append(payload.encode())
        elif kind == "replace":
            end, new = payload
            out.append(new.encode())
            pos = end
        elif kind == "init":
            end = payload
            init = data[off:end].decode()
            out.append(("{" + ", ".join([init] * MAX_RANKS) + "}").encode())
            pos = end
        elif kind == "main":
            name_end = payload
            # main's name and its parameter list
            i = name_end
            while i < len(data) and data[i:i + 1] != b"(":
                i += 1
            depth, j = 0, i
            while j < len(data):
                if data[j:j + 1] == b"(":
                    depth += 1
                elif data[j:j + 1] == b")":
                    depth -= 1
                    if depth == 0:
                        break
                j += 1
            out.append(b"__gpu_main(int argc, char** argv)")
            pos = j + 1
    out.append(data[pos:])
    return b"".join(out).decode()


def defines_main(ast, src):
    lc = Locator()
    main_real = os.path.realpath(src)
    found = False

    def visit(n):
        nonlocal found
        loc = lc.see(n.get("loc"))
        rng = n.get("range", {})
        lc.see(rng.get("begin"))
        lc.see(rng.get("end"))
        if (n.get("kind") == "FunctionDecl" and n.get("name") == "main" and loc and loc[0]
                and os.path.realpath(loc[0]) == main_real and any(c.get("kind") == "CompoundStmt"
                                                                  for c in n.get("inner", []))):
            found = True
        for c in n.get("inner", []):
            visit(c)

    visit(ast)
    return found


def convert(src, dst, include_dirs=()):
    """Convert C source `src` into HIP source `dst`; returns whether it
    defines main (and so got the host entry)."""
    with open(src) as f:
        text = f.read()
    with tempfile.TemporaryDirectory() as tmp:
        ast = parse_ast(src, list(include_dirs), tmp)
        edits = collect(ast, src, os.path.join(tmp, "include"))
    has_main = defines_main(ast, src)
    body = apply(text, edits)
    header = (f"// Converted from {os.path.basename(src)} for the device MPI runtime (dphpc-gpu_mpi_amd/"
              f"converter.py).\n#include \"gpumpi/converted.h\"\n#line 1 \"{os.path.basename(src)}\"\n")
    tail = "\n#include \"gpumpi/gpu_main.h\"\n" if has_main else "\n"
    with open(dst, "w") as f:
        f.write(header + body + tail)
    return has_main


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("src")
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("-I", dest="include", action="append", default=[])
    a = ap.parse_args(argv)
    convert(a.src, a.output, a.include)
    return 0


if __name__ == "__main__":
    sys.exit(main())
