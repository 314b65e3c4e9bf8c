"""Readers for the POV-Ray files CueMol writes: the scene file (.pov) for the
viewing setup and the geometry include (.inc) for meshes, spheres and
cylinders, and a loader that turns the pair into a Scene.

Neither is a POV-Ray interpreter. Both evaluate the small part of the scene
description language CueMol's templates use:

  read_pov_scene     src/bench/pov/pov_scene_reader.{hpp,cpp}: #declare /
                     #local constants, #if/#ifdef/#ifndef/#else/#end, skipped
                     #macro bodies, camera (orthographic | perspective),
                     background, SpecLighting/FlashLighting calls, fog,
                     global_settings assumed_gamma, the `_scene` include path
  read_geometry      src/bench/geom/mesh2_*.cpp: mesh2 blocks welded into one
                     indexed mesh, sphere{}, cylinder{} (capped unless `open`),
                     edge_line/edge_line2 calls (open cylinders tagged as
                     edges), declared textures and finishes, `_show_<id>`
                     sections as transparency groups
  load_pov_scene     src/bench/scene_setup.cpp:26-230 (buildSceneFromPov)
  to_principled_material  src/bench/material_convert.cpp (--material principled)

Values are double-precision vectors of up to four components (a scalar has
one and broadcasts); unknown identifiers and functions evaluate to an empty
value (reads as zero), as the reference tolerates them.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field, replace
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .scene import Camera, DistantLight, Fog, Material, Scene

Vec3 = Tuple[float, float, float]

# --------------------------------------------------------------------------
# tokens  (src/bench/pov/pov_lexer.cpp:9-95)
# --------------------------------------------------------------------------
NUM, IDENT, STR, PUNCT, END = range(5)


@dataclass
class Token:
    kind: int
    s: str = ""
    num: float = 0.0
    line: int = 1


def tokenize(src: str) -> List[Token]:
    """Numbers (digits, optional fraction and exponent; a leading '.' needs a
    digit after it), identifiers ([A-Za-z_][A-Za-z0-9_]*), double-quoted
    strings, // and /* */ comments, and every other character as a one-
    character punctuation token; a final END token."""
    out: List[Token] = []
    n, i, line = len(src), 0, 1
    while i < n:
        c = src[i]
        if c == "\n":
            line += 1
            i += 1
            continue
        if c.isspace():
            i += 1
            continue
        if c == "/" and i + 1 < n and src[i + 1] == "/":
            while i < n and src[i] != "\n":
                i += 1
            continue
        if c == "/" and i + 1 < n and src[i + 1] == "*":
            j = src.find("*/", i + 2)
            stop = n if j < 0 else j + 2
            line += src.count("\n", i, stop)
            i = stop
            continue
        if c == '"':
            j = src.find('"', i + 1)
            stop = n if j < 0 else j
            out.append(Token(STR, src[i + 1:stop], 0.0, line))
            line += src.count("\n", i, stop)
            i = stop + 1 if j >= 0 else n
            continue
        if c.isdigit() and c.isascii() or (c == "." and i + 1 < n and src[i + 1].isdigit()):
            j = i
            while j < n and src[j].isdigit():
                j += 1
            if j < n and src[j] == ".":
                j += 1
                while j < n and src[j].isdigit():
                    j += 1
            if j < n and src[j] in "eE":
                k = j + 1
                if k < n and src[k] in "+-":
                    k += 1
                if k < n and src[k].isdigit():
                    j = k
                    while j < n and src[j].isdigit():
                        j += 1
            out.append(Token(NUM, "", float(src[i:j]), line))
            i = j
            continue
        if c.isalpha() and c.isascii() or c == "_":
            j = i
            while j < n and (src[j].isalnum() and src[j].isascii() or src[j] == "_"):
                j += 1
            out.append(Token(IDENT, src[i:j], 0.0, line))
            i = j
            continue
        out.append(Token(PUNCT, c, 0.0, line))
        i += 1
    out.append(Token(END, "", 0.0, line))
    return out


# --------------------------------------------------------------------------
# values  (src/bench/pov/pov_lexer.cpp:97-114)
# --------------------------------------------------------------------------
Value = Tuple[float, ...]  # 0..4 components


def comp(v: Value, i: int) -> float:
    """Component i; a scalar broadcasts, a missing component is 0."""
    if not v:
        return 0.0
    if len(v) == 1:
        return v[0]
    return v[i] if i < len(v) else 0.0


def bin_op(a: Value, b: Value, op: str) -> Value:
    """Componentwise arithmetic; scalar op scalar is a scalar, otherwise the
    wider operand's width; division by zero gives 0."""
    n = 1 if (len(a) <= 1 and len(b) <= 1) else max(len(a), len(b), 1)
    r = []
    for i in range(n):
        x, y = comp(a, i), comp(b, i)
        if op == "+":
            r.append(x + y)
        elif op == "-":
            r.append(x - y)
        elif op == "*":
            r.append(x * y)
        else:
            r.append(x / y if y != 0.0 else 0.0)
    return tuple(r)


def _f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0]


def _vec3(v: Value) -> Vec3:
    return (_f32(comp(v, 0)), _f32(comp(v, 1)), _f32(comp(v, 2)))


def _normalize(v) -> Vec3:
    a = np.asarray(v, np.float32)
    return tuple(float(x) for x in a / np.float32(np.sqrt(np.float32(np.dot(a, a)))))


def _length(v) -> float:
    a = np.asarray(v, np.float32)
    return float(np.sqrt(np.float32(np.dot(a, a))))


_COLOR_KW = {"rgb", "rgbf", "rgbt", "rgbft", "srgb", "srgbf", "srgbt", "srgbft"}


class PovError(RuntimeError):
    """The input cannot be read as a CueMol POV-Ray file."""


class _Cursor:
    """Token cursor, brace matching and the expression grammar shared by the
    two readers: sum := term (('+'|'-') term)*, term := primary (('*'|'/')
    primary)*, primary := unary -/+ | '(' sum ')' | '<' sum, ... '>' (at most
    four components, each the first component of its sum) | number |
    identifier [call]."""

    def __init__(self, toks: List[Token]):
        self.t = toks
        self.pos = 0
        self.values: Dict[str, Value] = {}

    def cur(self) -> Token:
        return self.t[self.pos]

    def at_end(self) -> bool:
        return self.t[self.pos].kind == END

    def advance(self) -> None:
        if not self.at_end():
            self.pos += 1

    def is_punct(self, p: str) -> bool:
        c = self.t[self.pos]
        return c.kind == PUNCT and c.s == p

    def is_ident(self, name: Optional[str] = None) -> bool:
        c = self.t[self.pos]
        return c.kind == IDENT and (name is None or c.s == name)

    def punct_at(self, k: int, p: str) -> bool:
        return k < len(self.t) and self.t[k].kind == PUNCT and self.t[k].s == p

    def match_brace(self, open_idx: int) -> int:
        depth = 0
        for k in range(open_idx, len(self.t)):
            tk = self.t[k]
            if tk.kind == PUNCT and tk.s == "{":
                depth += 1
            elif tk.kind == PUNCT and tk.s == "}":
                depth -= 1
                if depth == 0:
                    return k
            elif tk.kind == END:
                break
        raise PovError(f"unbalanced '{{' near line {self.t[open_idx].line}")

    # expressions
    def call(self, name: str) -> Value:
        """A call name(...) at the cursor (on the '('); the geometry reader
        skips calls, the scene reader evaluates a few functions."""
        depth = 0
        while True:
            if self.is_punct("("):
                depth += 1
            elif self.is_punct(")"):
                depth -= 1
            self.advance()
            if depth <= 0 or self.at_end():
                return ()

    def primary(self) -> Value:
        if self.is_punct("-"):
            self.advance()
            return tuple(-x for x in self.primary())
        if self.is_punct("+"):
            self.advance()
            return self.primary()
        if self.is_punct("("):
            self.advance()
            v = self.sum()
            if self.is_punct(")"):
                self.advance()
            return v
        if self.is_punct("<"):
            self.advance()
            r: List[float] = []
            while not self.is_punct(">") and not self.at_end() and len(r) < 4:
                e = self.sum()
                r.append(e[0] if e else 0.0)
                if self.is_punct(","):
                    self.advance()
                else:
                    break
            if self.is_punct(">"):
                self.advance()
            return tuple(r)
        c = self.cur()
        if c.kind == NUM:
            self.advance()
            return (c.num,)
        if c.kind == IDENT:
            self.advance()
            if self.is_punct("("):
                return self.call(c.s)
            return self.values.get(c.s, ())
        return ()

    def term(self) -> Value:
        a = self.primary()
        while self.is_punct("*") or self.is_punct("/"):
            op = self.cur().s
            self.advance()
            a = bin_op(a, self.primary(), op)
        return a

    def sum(self) -> Value:
        a = self.term()
        while self.is_punct("+") or self.is_punct("-"):
            op = self.cur().s
            self.advance()
            a = bin_op(a, self.term(), op)
        return a

    def call_args(self) -> List[Value]:
        args: List[Value] = []
        if not self.is_punct("("):
            return args
        self.advance()
        if not self.is_punct(")"):
            args.append(self.sum())
            while self.is_punct(","):
                self.advance()
                args.append(self.sum())
        if self.is_punct(")"):
            self.advance()
        return args

    def skip_macro(self) -> None:
        """From 'macro' past its matching #end; nested #macro/#if/#ifdef/
        #ifndef/#while/#for/#switch open a level."""
        self.advance()
        depth = 1
        while not self.at_end() and depth > 0:
            if self.is_punct("#"):
                self.advance()
                if self.is_ident():
                    d = self.cur().s
                    if d in ("macro", "if", "ifdef", "ifndef", "while", "for", "switch"):
                        depth += 1
                    elif d == "end":
                        depth -= 1
                    self.advance()
            else:
                self.advance()

    def skip_to_semicolon(self) -> None:
        """Past the next ';' outside brackets, or up to the next '#'."""
        depth = 0
        while not self.at_end():
            c = self.cur()
            if c.kind == PUNCT:
                if c.s in "{([":
                    depth += 1
                elif c.s in "})]":
                    depth = max(0, depth - 1)
                elif depth == 0 and c.s == ";":
                    self.advance()
                    return
                elif depth == 0 and c.s == "#":
                    return
            self.advance()


# --------------------------------------------------------------------------
# scene file  (src/bench/pov/pov_scene_reader.cpp)
# --------------------------------------------------------------------------
@dataclass
class PovFog:
    """fog { ... } as read: the POV fields plus the linear-fog start/end
    derived from the camera distance (scene.hpp Fog)."""
    enabled: bool = False
    distance: float = 0.0
    type: int = 1
    offset: float = 0.0
    alt: float = 0.0
    up: Vec3 = (0.0, 1.0, 0.0)
    color: Vec3 = (0.0, 0.0, 0.0)
    start: float = 0.0
    end: float = 0.0

    def to_fog(self) -> Fog:
        return Fog(enabled=self.enabled, color=self.color, start=self.start, end=self.end)


@dataclass
class PovSceneResult:
    """PovSceneResult (pov_scene_reader.hpp:45-62)."""
    camera: Camera = field(default_factory=Camera)
    lights: List[DistantLight] = field(default_factory=list)
    background: Vec3 = (0.0, 0.0, 0.0)
    fog: PovFog = field(default_factory=PovFog)
    assumed_gamma: float = 1.0
    declares: Dict[str, List[float]] = field(default_factory=dict)
    include_path: str = ""
    radiosity: bool = False


_FUNCS = {
    "degrees": lambda s: s(0) * 57.295779513082323,
    "radians": lambda s: s(0) * 0.017453292519943295,
    "atan2": lambda s: math.atan2(s(0), s(1)),
    "atan": lambda s: math.atan(s(0)),
    "tan": lambda s: math.tan(s(0)),
    "sin": lambda s: math.sin(s(0)),
    "cos": lambda s: math.cos(s(0)),
    "sqrt": lambda s: math.sqrt(max(0.0, s(0))),
    "abs": lambda s: abs(s(0)),
    "floor": lambda s: float(math.floor(s(0))),
    "int": lambda s: float(math.trunc(s(0))),
    "min": lambda s: min(s(0), s(1)),
    "max": lambda s: max(s(0), s(1)),
    "pow": lambda s: math.pow(s(0), s(1)),
}


def _apply_func(name: str, a: List[Value]) -> Value:
    """The built-in functions (pov_scene_reader.cpp applyFunc); scalar
    arguments read component 0, unknown functions give an empty value."""
    def s(i: int) -> float:
        return comp(a[i], 0) if i < len(a) else 0.0

    if name in _FUNCS:
        try:
            return (float(_FUNCS[name](s)),)
        except (ValueError, OverflowError):
            return (math.nan,)
    if name == "vlength" and a:
        v = a[0]
        return (math.sqrt(comp(v, 0) ** 2 + comp(v, 1) ** 2 + comp(v, 2) ** 2),)
    if name == "vnormalize" and a:
        v = a[0]
        ln = math.sqrt(comp(v, 0) ** 2 + comp(v, 1) ** 2 + comp(v, 2) ** 2)
        return tuple(comp(v, i) / ln for i in range(3)) if ln > 0.0 else (0.0, 0.0, 0.0)
    if name == "vdot" and len(a) >= 2:
        return (sum(comp(a[0], i) * comp(a[1], i) for i in range(3)),)
    if name == "vcross" and len(a) >= 2:
        x, y = a[0], a[1]
        return (comp(x, 1) * comp(y, 2) - comp(x, 2) * comp(y, 1),
                comp(x, 2) * comp(y, 0) - comp(x, 0) * comp(y, 2),
                comp(x, 0) * comp(y, 1) - comp(x, 1) * comp(y, 0))
    return ()


class _SceneReader(_Cursor):
    def __init__(self, toks: List[Token], width: int, height: int, predefined: Dict[str, float]):
        super().__init__(toks)
        self.values["image_width"] = (float(width),)
        self.values["image_height"] = (float(height),)
        for k, v in predefined.items():
            self.values[k] = (float(v),)
        self.stack: List[List[bool]] = []  # [parent active, condition, in else]
        self.in_camera = False
        self.cam_end = 0
        self.cam: Dict[str, object] = {}
        self.r = PovSceneResult()

    def call(self, name: str) -> Value:
        return _apply_func(name, self.call_args())

    def active(self) -> bool:
        if not self.stack:
            return True
        parent, cond, in_else = self.stack[-1]
        return parent and (not cond if in_else else cond)

    def run(self) -> PovSceneResult:
        while not self.at_end():
            if self.in_camera and self.pos >= self.cam_end:
                self.finish_camera()
                self.in_camera = False
                self.advance()
                continue
            if self.is_punct("#"):
                self.advance()
                self.directive()
            elif self.active() and self.is_ident():
                self.ident()
            else:
                self.advance()
        r = self.r
        if r.fog.enabled:
            if "_distance" in self.values:
                dist = comp(self.values["_distance"], 0)
            elif "location" in self.cam and "look_at" in self.cam:
                d = np.subtract(_vec3(self.cam["look_at"]), _vec3(self.cam["location"]))
                dist = _length(d)
            else:
                dist = 200.0
            dist = max(dist, 1.0)
            r.fog.start = _f32(dist)
            r.fog.end = _f32(dist + 1.5 * r.fog.distance)
        for k in sorted(self.values):
            v = self.values[k]
            r.declares[k] = list(v) if v else [0.0]
        r.radiosity = "_radiosity" in self.values
        return r

    def directive(self) -> None:
        if not self.is_ident():
            return
        d = self.cur().s
        if d == "macro":
            self.skip_macro()
        elif d in ("declare", "local"):
            self.advance()
            self.declare()
        elif d in ("ifdef", "ifndef"):
            self.advance()
            name = ""
            if self.is_punct("("):
                self.advance()
            if self.is_ident():
                name = self.cur().s
                self.advance()
            if self.is_punct(")"):
                self.advance()
            defined = name in self.values
            self.stack.append([self.active(), defined if d == "ifdef" else not defined, False])
        elif d in ("if", "while", "switch"):
            self.advance()
            v = self.primary() if self.is_punct("(") else self.sum()
            self.stack.append([self.active(), comp(v, 0) != 0.0, False])
        elif d in ("else", "elseif"):
            if self.stack:
                self.stack[-1][2] = True
            self.advance()
            if d == "elseif" and self.is_punct("("):
                self.primary()
        elif d in ("end", "break"):
            if d == "end" and self.stack:
                self.stack.pop()
            self.advance()
        elif d == "version":
            self.advance()
            self.skip_to_semicolon()
        elif d == "include":
            self.advance()
            if self.cur().kind == STR:
                self.advance()
        else:
            self.advance()

    def declare(self) -> None:
        if not self.is_ident():
            return
        name = self.cur().s
        self.advance()
        if not self.is_punct("="):
            return
        self.advance()
        if self.is_punct("#"):
            self.advance()
            if self.is_ident("include"):
                self.advance()
                if self.cur().kind == STR:
                    if self.active() and name == "_scene":
                        self.r.include_path = self.cur().s
                    self.advance()
            return
        if self.is_ident() and self.cur().s in ("texture", "pigment", "finish", "normal", "material"):
            self.advance()
            if self.is_punct("{"):
                self.pos = self.match_brace(self.pos) + 1
            return
        v = self.sum()
        if self.active() and v:
            self.values[name] = v
        self.skip_to_semicolon()

    def ident(self) -> None:
        name = self.cur().s
        if self.in_camera:
            self.camera_keyword(name)
            return
        self.advance()
        if name == "camera":
            if self.is_punct("{"):
                self.cam_end = self.match_brace(self.pos)
                self.advance()
                self.in_camera = True
                self.cam = {"direction": (0.0, 0.0, -1.0)}
            return
        if name == "background":
            self.background()
        elif name == "SpecLighting":
            self.spec_lighting(self.call_args())
        elif name == "FlashLighting":
            self.flash_lighting(self.call_args())
        elif name == "fog":
            self.fog()
        elif name == "global_settings":
            self.global_settings()
        elif self.is_punct("{"):
            self.pos = self.match_brace(self.pos) + 1
        elif self.is_punct("("):
            _Cursor.call(self, name)

    def camera_keyword(self, name: str) -> None:
        self.advance()
        if name in ("orthographic", "perspective"):
            self.cam[name] = True
        elif name in ("direction", "up", "right", "location", "look_at"):
            self.cam[name] = self.sum()
        elif name == "sky":
            self.sum()
        elif name == "angle":
            self.cam["angle"] = comp(self.sum(), 0)

    def finish_camera(self) -> None:
        """pov_scene_reader.cpp finalizeCamera: orthographic height = |up|;
        perspective fovy = degrees(2 atan(tan(radians(angle)/2) |up|/|right|)),
        angle 40 when absent; float arithmetic as the reference does it."""
        c = self.cam
        loc = _vec3(c["location"]) if "location" in c else (0.0, 0.0, 1.0)
        if "look_at" in c:
            d = _normalize(np.subtract(np.float32(_vec3(c["look_at"])), np.float32(loc)))
        else:
            d = _normalize(_vec3(c["direction"]))
        up = _normalize(_vec3(c["up"])) if "up" in c else (0.0, 1.0, 0.0)
        cam = Camera(position=loc, direction=d, up=up)
        if c.get("orthographic"):
            cam.orthographic = True
            cam.height = _length(_vec3(c["up"])) if "up" in c else 1.0
        else:
            rl = _length(_vec3(c["right"])) if "right" in c else 1.0
            ul = _length(_vec3(c["up"])) if "up" in c else 1.0
            fovx = c.get("angle", 40.0)
            half = np.float32(np.float32(fovx) * np.float32(math.pi / 180.0)) * np.float32(0.5)
            t = float(np.tan(half, dtype=np.float32)) * (ul / rl)
            a = np.float32(np.float32(2.0) * np.float32(math.atan(t)))
            cam.fovy = float(np.float32(a * np.float32(180.0 / math.pi)))
        self.r.camera = cam

    def _color_at(self, open_idx: int, close: int) -> Optional[Value]:
        """The first color expression in a block: `color [kind] expr`,
        `colour ...` or `kind expr` with kind one of rgb/rgbf/rgbt/rgbft and
        their srgb forms."""
        for k in range(open_idx + 1, close):
            tk = self.t[k]
            if tk.kind != IDENT:
                continue
            if tk.s in ("color", "colour"):
                p = k + 1
                if p < close and self.t[p].kind == IDENT and self.t[p].s in _COLOR_KW:
                    p += 1
            elif tk.s in _COLOR_KW:
                p = k + 1
            else:
                continue
            self.pos = p
            return self.sum()
        return None

    def background(self) -> None:
        if not self.is_punct("{"):
            return
        o = self.pos
        close = self.match_brace(o)
        v = self._color_at(o, close)
        if v is not None:
            self.r.background = _vec3(v)
        self.pos = close + 1

    def global_settings(self) -> None:
        if not self.is_punct("{"):
            return
        o = self.pos
        close = self.match_brace(o)
        for k in range(o + 1, close):
            if self.t[k].kind == IDENT and self.t[k].s == "assumed_gamma":
                self.pos = k + 1
                self.r.assumed_gamma = _f32(comp(self.sum(), 0))
                break
        self.pos = close + 1

    def fog(self) -> None:
        if not self.is_punct("{"):
            return
        o = self.pos
        close = self.match_brace(o)
        f = self.r.fog
        f.enabled = True
        self.pos = o + 1
        while self.pos < close and not self.at_end():
            if not self.is_ident():
                self.advance()
                continue
            key = self.cur().s
            self.advance()
            if key == "distance":
                f.distance = _f32(comp(self.sum(), 0))
            elif key == "fog_type":
                f.type = int(comp(self.sum(), 0))
            elif key == "fog_offset":
                f.offset = _f32(comp(self.sum(), 0))
            elif key == "fog_alt":
                f.alt = _f32(comp(self.sum(), 0))
            elif key == "up":
                f.up = _vec3(self.sum())
            elif key in ("color", "colour"):
                if self.is_ident() and self.cur().s in _COLOR_KW:
                    self.advance()
                f.color = _vec3(self.sum())
            elif key in _COLOR_KW:
                f.color = _vec3(self.sum())
        self.pos = close + 1

    def spec_lighting(self, a: List[Value]) -> None:
        """SpecLighting(spread, distance, intensity, shadow): a key light
        travelling along -<1,1,1>; it casts highlights when spread > 1 or
        shadow != 0, with angular radius atan(spread / 40) when spread > 1."""
        if len(a) < 3:
            return
        spread = comp(a[0], 0)
        shadow = comp(a[3], 0) if len(a) > 3 else 0.0
        frm = _normalize((1.0, 1.0, 1.0))
        L = DistantLight(direction=_normalize(np.subtract(np.float32(0.0), np.float32(frm))),
                         color=(1.0, 1.0, 1.0), intensity=_f32(comp(a[2], 0)),
                         casts_highlight=(spread > 1.0 or shadow != 0.0),
                         angular_radius=_f32(math.atan(spread / 40.0)) if spread > 1.0 else 0.0)
        self.r.lights.append(L)

    def flash_lighting(self, a: List[Value]) -> None:
        """FlashLighting(intensity): a headlight from the eye
        (_stereo * _distance * _iod, 0, _distance) towards the origin, no
        highlight."""
        if not a:
            return

        def sym(name: str, dflt: float) -> float:
            return comp(self.values[name], 0) if name in self.values else dflt

        dist = sym("_distance", 200.0)
        eye = (_f32(sym("_stereo", 0.0) * dist * sym("_iod", 0.03)), 0.0, _f32(dist))
        L = DistantLight(direction=_normalize(np.subtract(np.float32(0.0), np.float32(eye))),
                         color=(1.0, 1.0, 1.0), intensity=_f32(comp(a[0], 0)), casts_highlight=False)
        self.r.lights.append(L)


def read_pov_scene_string(text: str, width: int = 300, height: int = 300,
                          predefined: Optional[Dict[str, float]] = None) -> PovSceneResult:
    """readPovSceneFromString (pov_scene_reader.hpp:68-69). predefined seeds
    the symbol table before the file is read (POV `Declare=name=value`), so
    an #ifndef default in the file does not override it."""
    return _SceneReader(tokenize(text), width, height, dict(predefined or {})).run()


def read_pov_scene(path: str, width: int = 300, height: int = 300,
                   predefined: Optional[Dict[str, float]] = None) -> PovSceneResult:
    """readPovScene (pov_scene_reader.hpp:72-73)."""
    try:
        with open(path, "rb") as f:
            text = f.read().decode("latin-1")
    except OSError as e:
        raise PovError(f"cannot open file: {path}") from e
    return read_pov_scene_string(text, width, height, predefined)


# --------------------------------------------------------------------------
# geometry include  (src/bench/geom/mesh2_*.cpp)
# --------------------------------------------------------------------------
# F_Metal* finishes of POV-Ray's metals.inc (ambient, brilliance, diffuse,
# specular, roughness, reflection), as mesh2_texture_parse.cpp:22-45 maps them
_NAMED_FINISH = {
    "F_MetalA": (0.35, 2.0, 0.3, 0.80, 1.0 / 20.0, 0.10),
    "F_MetalB": (0.30, 3.0, 0.4, 0.70, 1.0 / 60.0, 0.25),
    "F_MetalC": (0.25, 4.0, 0.5, 0.80, 1.0 / 80.0, 0.50),
    "F_MetalD": (0.15, 5.0, 0.6, 0.80, 1.0 / 100.0, 0.65),
    "F_MetalE": (0.10, 6.0, 0.7, 0.80, 1.0 / 120.0, 0.80),
}


def _srgb_to_linear(c: float) -> float:
    c = np.float32(c)
    if c <= np.float32(0.04045):
        return float(c / np.float32(12.92))
    return float(np.power((c + np.float32(0.055)) / np.float32(1.055), np.float32(2.4), dtype=np.float32))


@dataclass
class PovTexture:
    color: Tuple[float, float, float, float] = (1.0, 1.0, 1.0, 1.0)
    finish: Material = field(default_factory=Material)
    has_finish: bool = False


@dataclass
class GeomSphere:
    center: Vec3
    radius: float
    color: Tuple[float, float, float, float]
    material: Material
    group: int = 0
    from_edge: bool = False


@dataclass
class GeomCylinder:
    p0: Vec3
    p1: Vec3
    radius: float
    color: Tuple[float, float, float, float]
    material: Material
    group: int = 0
    opacity1: float = -1.0
    open: bool = False
    from_edge: bool = False


@dataclass
class SceneGeometry:
    """SceneGeometry (mesh2_reader.hpp:26-35): one indexed mesh (vertices
    split where position, normal or colour differ), its per-triangle block
    material and section group, the position classes of the 1e-4 weld grid,
    the analytic primitives, and the section names per group (index 0 is
    the default group, named '')."""
    positions: np.ndarray
    normals: np.ndarray
    colors: np.ndarray
    index: np.ndarray              # (T, 3) uint32
    pos_class: np.ndarray          # (V,) int32
    tri_material: np.ndarray       # (T,) uint8
    tri_group: np.ndarray          # (T,) uint16
    materials: List[Material]
    spheres: List[GeomSphere]
    cylinders: List[GeomCylinder]
    group_names: List[str]

    @property
    def n_triangles(self) -> int:
        return int(len(self.index))

    @property
    def material(self) -> Material:
        return self.materials[0] if self.materials else Material()


class _GeomReader(_Cursor):
    def __init__(self, toks: List[Token], symbols: Optional[Dict[str, Sequence[float]]]):
        super().__init__(toks)
        for k, v in (symbols or {}).items():
            self.values[k] = tuple(float(x) for x in list(v)[:4])
        self.textures: Dict[str, PovTexture] = {}
        self.verts: List[Vec3] = []
        self.norms: List[Vec3] = []
        self.tex_colors: List[Tuple[float, ...]] = []
        self.faces: List[Tuple[Tuple[int, int, int], Optional[Tuple[int, int, int]]]] = []
        self.out_pos: List[Tuple[float, ...]] = []
        self.out_nrm: List[Tuple[float, ...]] = []
        self.out_col: List[Tuple[float, ...]] = []
        self.out_cls: List[int] = []
        self.out_idx: List[int] = []
        self.tri_mat: List[int] = []
        self.tri_grp: List[int] = []
        self.class_map: Dict[Tuple[int, int, int], int] = {}
        self.block_materials: List[Material] = []
        self.block_material: Optional[Material] = None
        self.have_mesh = False
        self.spheres: List[GeomSphere] = []
        self.cylinders: List[GeomCylinder] = []
        self.group_names: List[str] = [""]
        self.group_index: Dict[str, int] = {}
        self.group = 0
        self.if_stack: List[int] = []

    def run(self) -> SceneGeometry:
        """mesh2_reader.cpp:22-70: directives, then mesh2 / sphere{ /
        cylinder{ / edge_line( / edge_line2( at the top level; everything
        else is stepped over token by token."""
        while not self.at_end():
            if self.is_punct("#"):
                self.advance()
                d = self.cur().s if self.is_ident() else ""
                if d == "macro":
                    self.skip_macro()
                elif d in ("declare", "local"):
                    self.advance()
                    self.declare()
                elif d in ("if", "ifdef", "ifndef"):
                    self.advance()
                    self.if_stack.append(self.group)
                    show = self.section_var()
                    if show:
                        self.group = self.group_for(show)
                elif d == "end":
                    self.advance()
                    if self.if_stack:
                        self.group = self.if_stack.pop()
                elif d:
                    self.advance()
            elif self.is_ident("mesh2"):
                self.advance()
                self.mesh2()
            elif self.is_ident("sphere") and self.punct_at(self.pos + 1, "{"):
                self.advance()
                self.sphere()
            elif self.is_ident("cylinder") and self.punct_at(self.pos + 1, "{"):
                self.advance()
                self.cylinder()
            elif self.is_ident() and self.cur().s in ("edge_line", "edge_line2") and self.punct_at(self.pos + 1, "("):
                two = self.cur().s == "edge_line2"
                self.advance()
                self.edge_line(two)
            else:
                self.advance()
        return self.finish()

    # sections
    def section_var(self) -> str:
        """The first identifier starting with `_show` inside the #if's
        parenthesised condition, or ''."""
        if not self.is_punct("("):
            return ""
        depth = 0
        for k in range(self.pos, len(self.t)):
            tk = self.t[k]
            if tk.kind == END:
                break
            if tk.kind == PUNCT and tk.s == "(":
                depth += 1
            elif tk.kind == PUNCT and tk.s == ")":
                depth -= 1
                if depth == 0:
                    break
            elif tk.kind == IDENT and tk.s.startswith("_show"):
                return tk.s
        return ""

    def group_for(self, show: str) -> int:
        sid = show[5:] if len(show) > 5 else show
        if sid not in self.group_index:
            self.group_index[sid] = len(self.group_names)
            self.group_names.append(sid)
        return self.group_index[sid]

    def declare(self) -> None:
        if not self.is_ident():
            return
        name = self.cur().s
        self.advance()
        if not self.is_punct("="):
            return
        self.advance()
        if self.is_ident("texture"):
            self.advance()
            if not self.is_punct("{"):
                return
            o = self.pos
            close = self.match_brace(o)
            tex = self.texture_body(o + 1, close)
            self.textures[name] = tex
            self.adopt(tex)
            self.pos = close + 1
            return
        if self.is_ident() and self.cur().s in ("pigment", "finish", "normal", "material"):
            self.advance()
            if self.is_punct("{"):
                self.pos = self.match_brace(self.pos) + 1
            return
        v = self.sum()
        if v:
            self.values[name] = v
        self.skip_to_semicolon()

    # textures  (mesh2_texture_parse.cpp)
    def adopt(self, tex: PovTexture) -> None:
        """The first finish seen since the last mesh2 block is that block's
        material."""
        if tex.has_finish and self.block_material is None:
            self.block_material = tex.finish

    def texture_body(self, begin: int, end: int) -> PovTexture:
        tex = PovTexture()
        have_color = False
        base = ""
        k = begin
        while k < end:
            tk = self.t[k]
            if tk.kind == IDENT:
                if tk.s in ("pigment", "finish") and k + 1 < end and self.punct_at(k + 1, "{"):
                    c = self.match_brace(k + 1)
                    if tk.s == "pigment":
                        col = self.pigment_color(k + 2, c)
                        if col is not None:
                            tex.color = col
                            have_color = True
                    else:
                        tex.finish = self.finish_body(k + 2, c)
                        tex.has_finish = True
                    k = c + 1
                    continue
                if not base and not have_color and tk.s not in ("pigment", "finish"):
                    base = tk.s  # a previously declared texture
                k += 1
            elif tk.kind == PUNCT and tk.s == "{":
                k = self.match_brace(k) + 1
            else:
                k += 1
        if base in self.textures:
            b = self.textures[base]
            if not have_color:
                tex.color = b.color
            if not tex.has_finish and b.has_finish:
                tex.finish = b.finish
                tex.has_finish = True
        return tex

    def pigment_color(self, begin: int, end: int):
        for k in range(begin, end):
            tk = self.t[k]
            if tk.kind != IDENT:
                continue
            if tk.s in ("color", "colour"):
                kind, p = "rgb", k + 1
                if p < end and self.t[p].kind == IDENT and self.t[p].s in _COLOR_KW:
                    kind = self.t[p].s
                    p += 1
            elif tk.s in _COLOR_KW:
                kind, p = tk.s, k + 1
            else:
                continue
            self.pos = p
            return self.to_color(self.sum(), kind)
        return None

    @staticmethod
    def to_color(v: Value, kind: str) -> Tuple[float, float, float, float]:
        """A colour expression of the given kind: three components as rgb, one
        broadcast to grey, none white; the t/f kinds read component 4 as
        transparency (opacity 1 - t); srgb kinds are decoded to linear."""
        if len(v) >= 3:
            r, g, b = _f32(v[0]), _f32(v[1]), _f32(v[2])
        elif len(v) >= 1:
            r = g = b = _f32(v[0])
        else:
            r = g = b = 1.0
        opacity = 1.0
        if kind.endswith(("t", "f")) and len(v) >= 4:
            opacity = float(np.float32(1.0) - np.float32(v[3]))
        if kind.startswith("srgb"):
            r, g, b = _srgb_to_linear(r), _srgb_to_linear(g), _srgb_to_linear(b)
        return (r, g, b, opacity)

    def finish_body(self, begin: int, end: int) -> Material:
        m = Material()
        k = begin
        keys = {"ambient": "ambient", "diffuse": "diffuse", "specular": "specular", "roughness": "roughness",
                "brilliance": "brilliance", "phong": "phong", "phong_size": "phong_size",
                "reflection": "reflection", "emission": "emission"}
        while k < end:
            tk = self.t[k]
            if tk.kind == PUNCT and tk.s == "{":
                k = self.match_brace(k) + 1  # nested block, e.g. reflection { ... }
                continue
            if tk.kind == IDENT and tk.s in keys:
                self.pos = k + 1
                v = self.sum()
                m = replace(m, **{keys[tk.s]: _f32(v[0]) if v else 0.0})
                k = self.pos
                continue
            if tk.kind == IDENT and tk.s == "metallic":
                n = k + 1
                m = replace(m, metallic=True)
                if n < end and (self.t[n].kind == NUM or self.punct_at(n, "-") or self.punct_at(n, "+")):
                    self.pos = n
                    v = self.sum()
                    m = replace(m, metallic=(_f32(v[0]) if v else 1.0) > 0.0)
                    k = self.pos
                else:
                    k = n
                continue
            if tk.kind == IDENT and tk.s in _NAMED_FINISH:
                a, br, d, s, r, refl = _NAMED_FINISH[tk.s]
                m = Material(ambient=_f32(a), brilliance=_f32(br), diffuse=_f32(d), specular=_f32(s),
                             roughness=_f32(r), reflection=_f32(refl), metallic=True, phong=0.0)
            k += 1
        return m

    def texture_in(self, begin: int, end: int) -> PovTexture:
        """The first texture { } between begin and end; black when absent."""
        for k in range(begin, end):
            if self.t[k].kind == IDENT and self.t[k].s == "texture" and k + 1 < end and self.punct_at(k + 1, "{"):
                return self.texture_body(k + 2, self.match_brace(k + 1))
        return PovTexture(color=(0.0, 0.0, 0.0, 1.0))

    # geometry blocks  (mesh2_block_parser.cpp)
    def mesh2(self) -> None:
        if not self.is_punct("{"):
            return
        o = self.pos
        close = self.match_brace(o)
        k = o + 1
        while k < close:
            tk = self.t[k]
            if tk.kind == IDENT and k + 1 < close and self.punct_at(k + 1, "{"):
                c = self.match_brace(k + 1)
                if tk.s == "vertex_vectors":
                    self.vec3_array(k + 2, c, self.verts)
                elif tk.s == "normal_vectors":
                    self.vec3_array(k + 2, c, self.norms)
                elif tk.s == "texture_list":
                    self.texture_list(k + 2, c)
                elif tk.s == "face_indices":
                    self.face_indices(k + 2, c)
                k = c + 1
                continue
            if tk.kind == PUNCT and tk.s == "{":
                k = self.match_brace(k) + 1
                continue
            k += 1
        self.have_mesh = True
        self.build_block()
        self.block_material = None
        self.pos = close + 1

    def signed_number(self) -> float:
        sign = 1.0
        while self.is_punct("-") or self.is_punct("+"):
            if self.is_punct("-"):
                sign = -sign
            self.advance()
        if self.cur().kind == NUM:
            x = _f32(self.cur().num)
            self.advance()
            return sign * x
        return 0.0

    def vec3_literal(self) -> Vec3:
        if not self.is_punct("<"):
            return (0.0, 0.0, 0.0)
        self.advance()
        c = [0.0, 0.0, 0.0]
        for i in range(3):
            c[i] = self.signed_number()
            if self.is_punct(","):
                self.advance()
            else:
                break
        while not self.is_punct(">") and not self.at_end():
            self.advance()
        if self.is_punct(">"):
            self.advance()
        return (c[0], c[1], c[2])

    def vec3_array(self, begin: int, end: int, dst: List[Vec3]) -> None:
        """`{ [count ,] <x,y,z>, ... }` of plain signed numbers."""
        self.pos = begin
        if self.cur().kind == NUM:
            save = self.pos
            self.advance()
            if self.is_punct(","):
                self.advance()
            else:
                self.pos = save
        while self.pos < end and not self.at_end():
            if self.is_punct("<"):
                dst.append(self.vec3_literal())
                if self.is_punct(","):
                    self.advance()
            elif self.is_punct(","):
                self.advance()
            else:
                break

    def texture_list(self, begin: int, end: int) -> None:
        k = begin
        if k < end and self.t[k].kind == NUM:
            k += 1
            if k < end and self.punct_at(k, ","):
                k += 1
        while k < end:
            if self.t[k].kind == IDENT and self.t[k].s == "texture" and k + 1 < end and self.punct_at(k + 1, "{"):
                c = self.match_brace(k + 1)
                tex = self.texture_body(k + 2, c)
                self.tex_colors.append(tex.color)
                self.adopt(tex)
                k = c + 1
                if k < end and self.punct_at(k, ","):
                    k += 1
                continue
            k += 1

    def face_indices(self, begin: int, end: int) -> None:
        """`{ [count ,] <a,b,c> [, t | , t0, t1, t2], ... }`: one texture id
        for the whole face or one per corner."""
        self.pos = begin
        if self.cur().kind == NUM:
            self.advance()
            if self.is_punct(","):
                self.advance()
        while self.pos < end and not self.at_end():
            if self.is_punct("<"):
                self.advance()
                v0 = int(self.signed_number())
                if self.is_punct(","):
                    self.advance()
                v1 = int(self.signed_number())
                if self.is_punct(","):
                    self.advance()
                v2 = int(self.signed_number())
                while not self.is_punct(">") and not self.at_end() and self.pos < end:
                    self.advance()
                if self.is_punct(">"):
                    self.advance()
                tex: List[int] = []
                while self.is_punct(",") and self.pos < end:
                    self.advance()
                    if self.cur().kind == NUM:
                        tex.append(int(self.cur().num))
                        self.advance()
                    else:
                        break
                t = None
                if len(tex) == 1:
                    t = (tex[0], tex[0], tex[0])
                elif len(tex) >= 3:
                    t = (tex[0], tex[1], tex[2])
                self.faces.append(((v0, v1, v2), t))
            elif self.is_punct(","):
                self.advance()
            else:
                break

    def sphere(self) -> None:
        if not self.is_punct("{"):
            return
        o = self.pos
        close = self.match_brace(o)
        self.pos = o + 1
        center = self.sum()
        if self.is_punct(","):
            self.advance()
        radius = _f32(comp(self.sum(), 0))
        tex = self.texture_in(self.pos, close)
        from_edge = any(self.t[k].kind == IDENT and self.t[k].s.endswith("_sl_tex") for k in range(o + 1, close))
        if radius > 0.0:
            self.spheres.append(GeomSphere(_vec3(center), radius, tex.color,
                                           tex.finish if tex.has_finish else Material.flat_outline(),
                                           self.group, from_edge))
        self.pos = close + 1

    def has_open(self, begin: int, end: int) -> bool:
        """`open` among the cylinder's own modifiers: at brace depth 0 and
        before its first texture/pigment/finish/normal/material."""
        depth = 0
        for k in range(begin, end):
            tk = self.t[k]
            if tk.kind == PUNCT and tk.s == "{":
                depth += 1
            elif tk.kind == PUNCT and tk.s == "}":
                depth = max(0, depth - 1)
            elif depth == 0 and tk.kind == IDENT:
                if tk.s == "open":
                    return True
                if tk.s in ("texture", "pigment", "finish", "normal", "material"):
                    break
        return False

    def cylinder(self) -> None:
        if not self.is_punct("{"):
            return
        o = self.pos
        close = self.match_brace(o)
        self.pos = o + 1
        p1 = self.sum()
        if self.is_punct(","):
            self.advance()
        p2 = self.sum()
        if self.is_punct(","):
            self.advance()
        radius = _f32(comp(self.sum(), 0))
        after = self.pos
        tex = self.texture_in(self.pos, close)
        if radius > 0.0:
            self.cylinders.append(GeomCylinder(_vec3(p1), _vec3(p2), radius, tex.color,
                                               tex.finish if tex.has_finish else Material.flat_outline(),
                                               self.group, open=self.has_open(after, close)))
        self.pos = close + 1

    def edge_line(self, two: bool) -> None:
        """edge_line(v1, n1, v2, n2, raise, w, tex, col) and edge_line2(v1, n1,
        t1, v2, n2, t2, raise, w, tex, col): an open cylinder from v1 + raise n1
        to v2 + raise n2 of radius w and colour col; edge_line2's t1/t2 are the
        end transparencies (opacity 1 - t clamped to [0, 1])."""
        a = self.call_args()
        if len(a) < (10 if two else 8):
            return
        iv2, in2, ir, iw, ic = (3, 4, 6, 7, 9) if two else (2, 3, 4, 5, 7)
        r = np.float32(comp(a[ir], 0))
        p0 = tuple(float(np.float32(x) + np.float32(n) * r) for x, n in zip(_vec3(a[0]), _vec3(a[1])))
        p1 = tuple(float(np.float32(x) + np.float32(n) * r) for x, n in zip(_vec3(a[iv2]), _vec3(a[in2])))
        col = self.to_color(a[ic], "rgb")
        op1 = -1.0
        if two:
            clamp = lambda t: float(min(np.float32(1.0), max(np.float32(0.0), np.float32(1.0) - np.float32(t))))
            col = (col[0], col[1], col[2], clamp(comp(a[2], 0)))
            op1 = clamp(comp(a[5], 0))
        w = _f32(comp(a[iw], 0))
        if w > 0.0:
            self.cylinders.append(GeomCylinder(p0, p1, w, col, Material.flat_outline(), self.group,
                                               op1, open=True, from_edge=True))

    # weld  (mesh2_build.cpp)
    def build_block(self) -> None:
        """De-index the block's faces into corners, give normal-less blocks
        flat face normals, then merge corners whose position, normal and
        colour are bit-identical (first-seen order); every new vertex gets
        the class of its 1e-4 weld cell, shared across blocks."""
        if not self.faces:
            return
        nv = len(self.verts)
        pos, nrm, col = [], [], []
        white = (1.0, 1.0, 1.0, 1.0)
        for vs, tex in self.faces:
            for k in range(3):
                vi = vs[k]
                if vi < 0 or vi >= nv:
                    raise PovError(f"face vertex index out of range: {vi}")
                pos.append(self.verts[vi])
                nrm.append(self.norms[vi] if vi < len(self.norms) else (0.0, 0.0, 0.0))
                c = white
                if tex is not None and self.tex_colors:
                    ti = min(max(tex[k], 0), len(self.tex_colors) - 1)
                    c = self.tex_colors[ti]
                col.append(c)
        P = np.asarray(pos, np.float32)
        N = np.asarray(nrm, np.float32)
        Ccol = np.asarray(col, np.float32)
        if not self.norms:
            a, b, c = P[0::3], P[1::3], P[2::3]
            n = np.cross(b - a, c - a).astype(np.float32)
            n = (n / np.sqrt((n * n).sum(1, dtype=np.float32))[:, None]).astype(np.float32)
            N = np.repeat(n, 3, axis=0)
        key = np.concatenate([P, N, Ccol], axis=1).view(np.uint32)
        vmap: Dict[bytes, int] = {}
        scaled = (P * np.float32(1.0e4)).astype(np.float64)  # float product, rounded half away from zero
        cells = np.floor(np.abs(scaled) + 0.5) * np.sign(scaled)
        for c in range(len(P)):
            kb = key[c].tobytes()
            vid = vmap.get(kb)
            if vid is None:
                vid = len(self.out_pos)
                vmap[kb] = vid
                self.out_pos.append(tuple(P[c]))
                self.out_nrm.append(tuple(N[c]))
                self.out_col.append(tuple(Ccol[c]))
                cell = tuple(int(x) for x in cells[c])
                cls = self.class_map.get(cell)
                if cls is None:
                    cls = len(self.class_map)
                    self.class_map[cell] = cls
                self.out_cls.append(cls)
            self.out_idx.append(vid)
        mat = len(self.block_materials)
        self.block_materials.append(self.block_material if self.block_material is not None else Material())
        self.tri_mat.extend([mat] * len(self.faces))
        self.tri_grp.extend([self.group] * len(self.faces))
        self.verts, self.norms, self.tex_colors, self.faces = [], [], [], []

    def finish(self) -> SceneGeometry:
        if not self.have_mesh and not self.spheres and not self.cylinders:
            raise PovError("no geometry (mesh2/sphere/edge_line) found in input")
        f3 = lambda a, w: np.asarray(a, np.float32).reshape(-1, w)
        return SceneGeometry(
            positions=f3(self.out_pos, 3), normals=f3(self.out_nrm, 3), colors=f3(self.out_col, 4),
            index=np.asarray(self.out_idx, np.uint32).reshape(-1, 3),
            pos_class=np.asarray(self.out_cls, np.int32),
            tri_material=np.asarray(self.tri_mat, np.uint8), tri_group=np.asarray(self.tri_grp, np.uint16),
            materials=self.block_materials, spheres=self.spheres, cylinders=self.cylinders,
            group_names=self.group_names)


def read_geometry_string(text: str, symbols: Optional[Dict[str, Sequence[float]]] = None) -> SceneGeometry:
    """readGeometryFromString (mesh2_reader.hpp:45-46); symbols are the
    constants the scene file declared (PovSceneResult.declares)."""
    return _GeomReader(tokenize(text), symbols).run()


def read_geometry(path: str, symbols: Optional[Dict[str, Sequence[float]]] = None) -> SceneGeometry:
    """readGeometryFromFile (mesh2_reader.hpp:49-50)."""
    try:
        with open(path, "rb") as f:
            text = f.read().decode("latin-1")
    except OSError as e:
        raise PovError(f"cannot open file: {path}") from e
    return read_geometry_string(text, symbols)


# --------------------------------------------------------------------------
# POV finish -> principled material  (src/bench/material_convert.cpp)
# --------------------------------------------------------------------------
def is_non_physical_finish(m: Material) -> bool:
    """Material::toonLike (include/umbreon/scene.hpp): toon model, or a POV
    finish whose look needs the POV model — highlight amount > 1,
    brilliance 0, or no lit term at all (ambient-only)."""
    if m.model == 2:
        return True
    if m.model != 0:
        return False
    if m.specular > 1.0 or m.phong > 1.0 or m.brilliance == 0.0:
        return True
    return m.diffuse <= 0.0 and m.specular <= 0.0 and m.phong <= 0.0 and m.reflection <= 0.0


def _exponent_to_roughness(e) -> float:
    """alpha = sqrt(2 / (e + 2)) of the Blinn exponent e, roughness =
    sqrt(alpha) (alpha clamped to 1); float arithmetic."""
    e = np.float32(e)
    alpha = np.sqrt(np.float32(2.0) / (e + np.float32(2.0)), dtype=np.float32)
    return float(np.sqrt(min(alpha, np.float32(1.0)), dtype=np.float32))


def to_principled_material(m: Material) -> Material:
    """toPrincipledMaterial (material_convert.cpp): a POV finish as the
    principled parameters — the Blinn highlight (exponent 1 / roughness,
    clamped to [1, 1e6]) or else the Phong one (exponent phong_size >= 1)
    gives the GGX roughness and the highlight amount the specular; a POV
    metallic finish is metallic 1; reflection without a highlight is a
    polished metal, with one the metallic amount is at least the reflection.
    Principled and non-physical (toon) materials come back unchanged."""
    if m.model == 1 or is_non_physical_finish(m):
        return m
    clamp01 = lambda x: float(min(np.float32(1.0), max(np.float32(0.0), np.float32(x))))
    high = m.specular > 0.0 or m.phong > 0.0
    if m.specular > 0.0:
        e = np.float32(1.0) / max(np.float32(m.roughness), np.float32(1e-6))
        e = min(max(e, np.float32(1.0)), np.float32(1.0e6))
        rough, spec = _exponent_to_roughness(e), clamp01(m.specular)
    elif m.phong > 0.0:
        rough, spec = _exponent_to_roughness(max(np.float32(1.0), np.float32(m.phong_size))), clamp01(m.phong)
    else:
        rough, spec = 0.5, 0.0
    metal = 1.0 if m.metallic else m.pbr_metallic
    if m.reflection > 0.0:
        if not high:
            metal, rough = 1.0, 0.0
        else:
            metal = max(metal, clamp01(m.reflection))
    return replace(m, model=1, pbr_roughness=rough, pbr_specular=spec, pbr_metallic=metal)


# --------------------------------------------------------------------------
# scene assembly  (src/bench/scene_setup.cpp:26-230)
# --------------------------------------------------------------------------
def geometry_to_scene(geo: SceneGeometry, scene: Optional[Scene] = None) -> Scene:
    """The geometry as Scene arrays: the mesh with its block materials and
    groups, and the primitives with their materials, groups, open flags and
    edge tags."""
    sc = scene if scene is not None else Scene()
    if geo.n_triangles:
        sc.set_mesh(geo.positions, geo.normals, geo.colors, geo.index, material=geo.material,
                    materials=geo.materials if len(geo.materials) > 1 else None,
                    tri_material=geo.tri_material if len(geo.materials) > 1 else None,
                    tri_group=geo.tri_group if geo.tri_group.any() else None)
    for s in geo.spheres:
        sc.add_sphere(s.center, s.radius, s.color, s.material, s.group, s.from_edge)
    for c in geo.cylinders:
        sc.add_cylinder(c.p0, c.p1, c.radius, c.color, c.material, c.group, c.opacity1, c.open, c.from_edge)
    return sc


def load_pov_scene(pov_path: str, width: int = 300, height: int = 300,
                   declares: Optional[Dict[str, float]] = None, gi: bool = False,
                   material_model: str = "pov") -> Tuple[Scene, List[str], float]:
    """A Scene from a CueMol .pov and the .inc it includes as `_scene`, the
    way the reference bench sets it up: with gi, the radiosity balance
    (_light_inten 1.6, _amb_frac 0.5, _flash_frac 0.5) replaces the defaults
    that were not declared otherwise and the ambient colour is
    _light_inten * _amb_frac (else white); assumed gamma 2.2. Returns the
    scene, the section names per group and the AO distance the bench uses
    (0.7 x the geometry's bounding diagonal, at least 1e-3). material_model
    "principled" converts every finish with to_principled_material except
    the edge decoration's (edge_line cylinders, `_sl_tex` spheres), as the
    bench's `--material principled` does."""
    if material_model not in ("pov", "principled"):
        raise ValueError(f"material_model must be 'pov' or 'principled', not {material_model!r}")
    decl = dict(declares or {})
    if gi:
        for k, plain, rad in (("_light_inten", 1.3, 1.6), ("_amb_frac", 0.0, 0.5), ("_flash_frac", 0.6, 0.5)):
            if k not in decl or decl[k] == plain:
                decl[k] = rad
    ps = read_pov_scene(pov_path, width, height, decl)
    if not ps.include_path:
        raise PovError(f"no '#declare _scene = #include \"...\"' found in {pov_path}")
    inc = ps.include_path if os.path.isabs(ps.include_path) else os.path.join(
        os.path.dirname(pov_path) or ".", ps.include_path)
    geo = read_geometry(inc, ps.declares)
    if material_model == "principled":
        geo.materials = [to_principled_material(m) for m in geo.materials]
        for p in geo.spheres + geo.cylinders:
            if not p.from_edge:
                p.material = to_principled_material(p.material)
    sc = geometry_to_scene(geo)
    lo = np.full(3, np.inf, np.float32)
    hi = np.full(3, -np.inf, np.float32)
    pts = [geo.positions]
    if geo.spheres:
        c = np.asarray([s.center for s in geo.spheres], np.float32)
        r = np.asarray([s.radius for s in geo.spheres], np.float32)[:, None]
        pts += [c - r, c + r]
    if geo.cylinders:
        pts += [np.asarray([c.p0 for c in geo.cylinders], np.float32), np.asarray([c.p1 for c in geo.cylinders], np.float32)]
    allp = np.concatenate([p.reshape(-1, 3) for p in pts])
    ao_distance = 0.0
    if len(allp):
        lo, hi = allp.min(0), allp.max(0)
        ao_distance = max(float(np.linalg.norm((hi - lo).astype(np.float32))) * 0.7, 1.0e-3)
    sc.camera = ps.camera
    sc.lights = list(ps.lights)
    sc.background = ps.background
    sc.fog = ps.fog.to_fog()
    if gi:
        a = _f32(decl["_light_inten"] * decl["_amb_frac"])
        sc.ambient_color = (a, a, a)
    else:
        sc.ambient_color = (1.0, 1.0, 1.0)
    sc.assumed_gamma = 2.2
    return sc, list(geo.group_names), ao_distance
