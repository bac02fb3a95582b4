"""Indented source text builder (upstream ``cutile/lowering/emitter.py:13-63``)."""

from __future__ import annotations

from contextlib import contextmanager


class CodeEmitter:
    def __init__(self, indent_str: str = "    "):
        self._lines: list = []
        self._level = 0
        self._indent = indent_str

    def line(self, text: str = "") -> None:
        self._lines.append(f"{self._indent * self._level}{text}" if text else "")

    def raw(self, text: str) -> None:
        """Append ``text`` verbatim (no indentation)."""
        self._lines.append(text)

    def blank(self) -> None:
        self._lines.append("")

    @contextmanager
    def indent(self):
        self._level += 1
        try:
            yield
        finally:
            self._level -= 1

    def render(self) -> str:
        return "\n".join(self._lines) + "\n"
