"""RON (Rusty Object Notation) reading and writing for the subset SyFrack's files use.

``to_ron`` writes dataclasses as named structs, dicts as anonymous structs, lists, tuples,
Option (None / Some), strings, booleans and numbers (floats by repr, so they read back exactly):
the form of ``state.ron``.  ``from_ron`` reads that subset and the reference's graph files
(graphs/*.ron: enum variants with fields, unit variants, tuples for glam vectors, line
comments) and scene files: a named struct or variant with fields becomes a dict with its name
under "__struct__", a variant with positional values (``Kelvin(4300.0)``, ``Graph((...))``) a
``Variant`` (a tuple that carries the name), a unit variant its name as a string, a tuple a tuple,
a list a list, Some(x) x.
"""
from __future__ import annotations

import dataclasses
import re
from typing import Any

_ESCAPE = re.compile(r"\\(u\{[0-9a-fA-F]+\}|x[0-9a-fA-F]{2}|.)", re.S)
_SIMPLE = {"n": "\n", "t": "\t", "r": "\r", "0": "\0", "\\": "\\", '"': '"', "'": "'"}


def _unescape(text: str) -> str:
    """RON string escapes: \\n \\t \\r \\0 \\\\ \\" \\' \\xNN \\u{N...}; other characters as they are."""
    def one(m):
        e = m.group(1)
        if e.startswith("u{"):
            return chr(int(e[2:-1], 16))
        if e.startswith("x"):
            return chr(int(e[1:], 16))
        return _SIMPLE.get(e, e)
    return _ESCAPE.sub(one, text)


def to_ron(value: Any, indent: int = 0) -> str:
    """Named structs for dataclasses, anonymous structs for dicts, lists, tuples, Option (None /
    Some), strings, booleans and numbers (floats with repr, so they read back exactly)."""
    pad, inner = "    " * indent, "    " * (indent + 1)
    if dataclasses.is_dataclass(value):
        fields = [(f.name, getattr(value, f.name)) for f in dataclasses.fields(value)]
        body = ",\n".join(f"{inner}{k}: {to_ron(v, indent + 1)}" for k, v in fields)
        return f"{type(value).__name__}(\n{body},\n{pad})"
    if isinstance(value, dict):
        if not value:
            return "()"
        body = ",\n".join(f"{inner}{k}: {to_ron(v, indent + 1)}" for k, v in value.items())
        return f"(\n{body},\n{pad})"
    if isinstance(value, bool):
        return "true" if value else "false"
    if value is None:
        return "None"
    if isinstance(value, int):
        return str(value)
    if isinstance(value, float):
        return "NaN" if value != value else repr(float(value))
    if isinstance(value, str):
        return '"' + value.replace("\\", "\\\\").replace('"', '\\"') + '"'
    if isinstance(value, tuple):
        return "(" + ", ".join(to_ron(v, indent + 1) for v in value) + ")"
    if isinstance(value, list):
        return "[" + ", ".join(to_ron(v, indent + 1) for v in value) + "]"
    raise TypeError(f"cannot write {type(value).__name__} as RON")


class Variant(tuple):
    """An enum variant with positional values: the values as a tuple, the name in ``variant``."""

    def __new__(cls, name: str, values):
        obj = super().__new__(cls, values)
        obj.variant = name
        return obj


_TOKEN = re.compile(r'\s*(?:(//[^\n]*)|("(?:[^"\\]|\\.)*")|([-+]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|inf(?![A-Za-z_0-9])|NaN(?![A-Za-z_0-9])))'
                    r'|(?:r#)?([A-Za-z_][A-Za-z_0-9]*)|(.))', re.S)


def from_ron(text: str) -> Any:
    """Parses the subset to_ron writes: named or anonymous structs become dicts (the name under
    "__struct__"), tuples tuples, lists lists, Some(x) x, None None."""
    toks = []
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        comment, string, number, ident, punct = m.groups()
        if comment is not None:
            continue
        if string is not None:
            toks.append(("str", _unescape(string[1:-1])))
        elif number is not None:
            toks.append(("num", float(number) if any(c in number for c in ".eEiN") else int(number)))
        elif ident is not None:
            toks.append(("id", ident))
        elif punct is not None and not punct.isspace():
            toks.append(("p", punct))
    i = 0

    def peek():
        return toks[i] if i < len(toks) else ("eof", None)

    def take(kind=None, val=None):
        nonlocal i
        t = peek()
        if (kind and t[0] != kind) or (val is not None and t[1] != val):
            raise ValueError(f"RON: expected {val or kind}, found {t[1]!r}")
        i += 1
        return t

    def value():
        t = peek()
        if t[0] in ("str", "num"):
            take()
            return t[1]
        if t[0] == "id":
            take()
            name = t[1]
            if name in ("true", "false"):
                return name == "true"
            if name == "None":
                return None
            if name == "Some":
                take("p", "(")
                v = value()
                take("p", ")")
                return v
            if peek() == ("p", "("):
                d = group()
                if isinstance(d, dict):
                    d["__struct__"] = name
                    return d
                return Variant(name, d if isinstance(d, tuple) else (d,))
            return name
        if t == ("p", "["):
            take()
            out = []
            while peek() != ("p", "]"):
                out.append(value())
                if peek() == ("p", ","):
                    take()
            take("p", "]")
            return out
        if t == ("p", "("):
            return group()
        raise ValueError(f"RON: unexpected {t[1]!r}")

    def group():
        take("p", "(")
        if peek()[0] == "id" and i + 1 < len(toks) and toks[i + 1] == ("p", ":"):
            out = {}
            while peek() != ("p", ")"):
                key = take("id")[1]
                take("p", ":")
                out[key] = value()
                if peek() == ("p", ","):
                    take()
            take("p", ")")
            return out
        items = []
        while peek() != ("p", ")"):
            items.append(value())
            if peek() == ("p", ","):
                take()
        take("p", ")")
        return tuple(items)

    v = value()
    if peek()[0] != "eof":
        raise ValueError(f"RON: trailing {peek()[1]!r}")
    return v
