"""Verilog-A preprocessor: `include, `define (object- and function-like, with
line continuations), `undef, `ifdef / `ifndef / `else / `elsif / `endif, and
macro expansion.  Comments are removed first; string literals are kept."""
from __future__ import annotations

import os
import re
from typing import Dict, List, Optional, Sequence, Tuple


class PreprocessError(Exception):
    pass


def strip_comments(text: str) -> str:
    out = []
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            if j < 0:
                break
            # a trailing backslash-continuation inside a comment of a macro
            # body keeps the line join
            seg = text[i:j]
            if seg.rstrip().endswith("\\"):
                out.append("\\")
            i = j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            if j < 0:
                raise PreprocessError("unterminated /* comment")
            out.append(" " * 0 + "\n" * text.count("\n", i, j))
            i = j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


_ident = re.compile(r"[A-Za-z_][A-Za-z0-9_$]*")


class Macro:
    def __init__(self, name: str, params: Optional[List[str]], body: str):
        self.name, self.params, self.body = name, params, body


def _split_args(text: str, i: int) -> Tuple[List[str], int]:
    """Parse '(a, b(c,d), e)' starting at text[i] == '('; return args and the
    index after the closing paren."""
    assert text[i] == "("
    depth, j, cur, args = 0, i, [], []
    while j < len(text):
        c = text[j]
        if c == "(":
            depth += 1
            if depth > 1:
                cur.append(c)
        elif c == ")":
            depth -= 1
            if depth == 0:
                args.append("".join(cur).strip())
                return args, j + 1
            cur.append(c)
        elif c == "," and depth == 1:
            args.append("".join(cur).strip())
            cur = []
        else:
            cur.append(c)
        j += 1
    raise PreprocessError("unbalanced macro arguments")


class Preprocessor:
    def __init__(self, include_dirs: Sequence[str] = (), defines: Dict[str, str] = None,
                 undefines: Sequence[str] = ()):
        self.include_dirs = list(include_dirs)
        self.macros: Dict[str, Macro] = {}
        for k, v in (defines or {}).items():
            self.macros[k] = Macro(k, None, v)
        self.blocked = set(undefines)   # never defined, whatever the source says

    def expand(self, text: str, depth: int = 0) -> str:
        if depth > 50:
            raise PreprocessError("macro expansion too deep")
        out, i, n = [], 0, len(text)
        changed = False
        while i < n:
            c = text[i]
            if c == '"':
                j = i + 1
                while j < n and text[j] != '"':
                    j += 1
                out.append(text[i:j + 1])
                i = j + 1
                continue
            if c == "`":
                m = _ident.match(text, i + 1)
                if m and m.group(0) in self.macros:
                    mac = self.macros[m.group(0)]
                    j = m.end()
                    if mac.params is not None:
                        k = j
                        while k < n and text[k] in " \t":
                            k += 1
                        if k >= n or text[k] != "(":
                            raise PreprocessError(f"macro `{mac.name} needs arguments")
                        args, j = _split_args(text, k)
                        if len(args) != len(mac.params):
                            raise PreprocessError(f"macro `{mac.name}: {len(args)} arguments")
                        body = mac.body
                        # textual substitution of whole-word parameters
                        body = _ident.sub(lambda mm: args[mac.params.index(mm.group(0))]
                                          if mm.group(0) in mac.params else mm.group(0), body)
                        out.append(" " + body + " ")
                    else:
                        out.append(" " + mac.body + " ")
                    i = j
                    changed = True
                    continue
            out.append(c)
            i += 1
        res = "".join(out)
        return self.expand(res, depth + 1) if changed else res

    def run_file(self, path: str) -> str:
        with open(path) as f:
            text = f.read()
        return self.run(text, os.path.dirname(os.path.abspath(path)))

    def _find(self, name: str, here: str) -> str:
        for d in [here] + self.include_dirs:
            p = os.path.join(d, name)
            if os.path.exists(p):
                return p
        raise PreprocessError(f"include file not found: {name}")

    def run(self, text: str, here: str) -> str:
        text = strip_comments(text)
        # join continuation lines
        text = re.sub(r"\\[ \t]*\n", " ", text)
        out: List[str] = []
        stack: List[Tuple[bool, bool]] = []   # (active, any branch taken)
        active = True
        for line in text.split("\n"):
            s = line.strip()
            if s.startswith("`ifdef") or s.startswith("`ifndef"):
                name = s.split()[1]
                cond = (name in self.macros) == s.startswith("`ifdef")
                stack.append((active, cond))
                active = active and cond
                continue
            if s.startswith("`elsif"):
                name = s.split()[1]
                parent, taken = stack[-1]
                cond = (not taken) and name in self.macros
                stack[-1] = (parent, taken or cond)
                active = parent and cond
                continue
            if s.startswith("`else"):
                parent, taken = stack[-1]
                stack[-1] = (parent, True)
                active = parent and not taken
                continue
            if s.startswith("`endif"):
                parent, _ = stack.pop()
                active = parent
                continue
            if not active:
                continue
            if s.startswith("`define"):
                rest = s[len("`define"):].strip()
                m = _ident.match(rest)
                name = m.group(0)
                params = None
                body = rest[m.end():]
                if body.startswith("("):
                    close = body.index(")")
                    params = [p.strip() for p in body[1:close].split(",") if p.strip()]
                    body = body[close + 1:]
                if name not in self.blocked:
                    self.macros[name] = Macro(name, params, body.strip())
                continue
            if s.startswith("`undef"):
                self.macros.pop(s.split()[1], None)
                continue
            if s.startswith("`include"):
                fname = s.split('"')[1]
                p = self._find(fname, here)
                with open(p) as f:
                    inc = f.read()
                out.append(self.run(inc, os.path.dirname(p)))
                continue
            if s.startswith("`timescale") or s.startswith("`default_nettype"):
                continue
            out.append(self.expand(line))
        if stack:
            raise PreprocessError("unterminated `ifdef")
        return "\n".join(out)


def preprocess(path: str, include_dirs: Sequence[str] = (), defines: Dict[str, str] = None,
               undefines: Sequence[str] = ()) -> str:
    return Preprocessor(include_dirs, defines, undefines).run_file(path)
