"""Static checks of the Python sources, above all the scripts that run only on a GPU (bench.py,
__graft_entry__.py, the GPU tests), so that a name error in a branch the CPU suite cannot execute
is caught here: every name a function reads must be a parameter, a name it binds, a name an
enclosing function binds, a module-level name or a builtin."""
import ast
import builtins
import glob
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bound(node):
    """Names a function (or comprehension scope) binds itself."""
    names = set()
    args = node.args
    for a in args.posonlyargs + args.args + args.kwonlyargs:
        names.add(a.arg)
    for a in (args.vararg, args.kwarg):
        if a is not None:
            names.add(a.arg)
    for sub in ast.walk(node):
        if isinstance(sub, ast.Name) and isinstance(sub.ctx, (ast.Store, ast.Del)):
            names.add(sub.id)
        elif isinstance(sub, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)) and sub is not node:
            names.add(sub.name)
        elif isinstance(sub, (ast.Import, ast.ImportFrom)):
            for a in sub.names:
                names.add((a.asname or a.name).split(".")[0])
        elif isinstance(sub, ast.ExceptHandler) and sub.name:
            names.add(sub.name)
        elif isinstance(sub, (ast.Global, ast.Nonlocal)):
            names.update(sub.names)
        elif isinstance(sub, ast.arg):
            names.add(sub.arg)
    return names


def _undefined(path):
    tree = ast.parse(open(path).read(), path)
    module = set(dir(builtins)) | {"__file__", "__name__"}
    for node in tree.body:
        if isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)):
            module.add(node.name)
        elif isinstance(node, (ast.Import, ast.ImportFrom)):
            for a in node.names:
                module.add((a.asname or a.name).split(".")[0])
        else:
            for sub in ast.walk(node):
                if isinstance(sub, ast.Name) and isinstance(sub.ctx, ast.Store):
                    module.add(sub.id)
    missing = []

    def visit(fn, outer):
        scope = outer | _bound(fn)
        for sub in ast.walk(fn):
            if isinstance(sub, ast.Name) and isinstance(sub.ctx, ast.Load) and sub.id not in scope:
                missing.append((fn.name, sub.id, sub.lineno))
        for sub in ast.walk(fn):
            if isinstance(sub, (ast.FunctionDef, ast.AsyncFunctionDef)) and sub is not fn:
                visit(sub, scope)

    for node in tree.body:
        if isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef)):
            visit(node, module)
    return missing


SOURCES = sorted(["bench.py", "__graft_entry__.py"] +
                 [os.path.relpath(p, ROOT) for d in ("nbounce_amd", "oracle", "tests", "tests/golden")
                  for p in glob.glob(os.path.join(ROOT, d, "*.py"))])


@pytest.mark.parametrize("name", SOURCES)
def test_sources_read_only_defined_names(name):
    missing = _undefined(os.path.join(ROOT, name))
    assert not missing, missing


def test_the_check_finds_an_undefined_name(tmp_path):
    p = tmp_path / "s.py"
    p.write_text("import os\nP = 1\ndef f(a):\n    b = a + P\n    return b + W * os.sep.count('x')\n")
    assert [m[1] for m in _undefined(str(p))] == ["W"]
