"""PyPSA networks as the LOPF builder reads them: the component schema and
PyPSA's CSV folder format (network-model's Schema.scala, CsvReader.scala,
ComponentTable.scala, Topology.scala).

A network folder holds `snapshots.csv` (labels plus the `objective`, `stores`
and `generators` weightings), one static file per component list
(`generators.csv`) and one file per time-varying attribute
(`generators-p_max_pu.csv`) covering only the entities that vary. Columns left
at their default are omitted, so every read falls through three levels: a
per-snapshot override, the entity's static value, the schema default
(ComponentTable.scala:117-127).

The schema (`schema.json`, PyPSA 1.2.4's component registry: 16 component
types, their attribute types, defaults and categories) ships with the package
beside PyPSA's standard line and transformer type library, because a CSV
export carries neither.
"""
from __future__ import annotations

import csv
import json
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "network_data")
_PORT = re.compile(r"bus\d+$")


class MalformedNetwork(RuntimeError):
    """CsvReader.MalformedNetwork (CsvReader.scala:32)."""


def is_port_column(name: str) -> bool:
    """Topology.isPortColumn (Topology.scala:95-99): `bus0`, `bus1`, `bus2`, ..."""
    return _PORT.match(name) is not None


# --------------------------------------------------------------------------
# schema


@dataclass
class Attribute:
    """AttributeSpec (Schema.scala:84-101): storage kind, default (already
    parsed to the kind) and whether PyPSA lets it vary per snapshot."""

    name: str
    kind: str  # "float" | "int" | "str" | "bool"
    default: object
    varying: bool


@dataclass
class ComponentSpec:
    name: str
    list_name: str
    category: str
    attributes: Dict[str, Attribute]

    def attribute(self, name: str) -> Optional[Attribute]:
        return self.attributes.get(name)


def _kind(pypsa_type: str) -> str:
    """AttributeType.parse (Schema.scala:19-36): PyPSA's `type` names either
    the element type or, for time-varying inputs, the shape ("static or
    series"), which is always numeric."""
    t = pypsa_type.strip().lower()
    for prefix, kind in (("float", "float"), ("int", "int"), ("string", "str"),
                         ("bool", "bool"), ("geometry", "str"), ("static", "float"),
                         ("series", "float")):
        if t.startswith(prefix):
            return kind
    raise ValueError(f"this reader cannot parse attributes of type '{pypsa_type}'")


def _default(raw, kind: str):
    """Schema.renderDefault (Schema.scala:176-191) plus the per-kind reading of
    ComponentTable (ComponentTable.scala:83-136): non-finite floats arrive
    tagged as {"$nan": true} or {"$inf": "+"|"-"}."""
    if isinstance(raw, dict):
        if "$nan" in raw:
            raw = math.nan
        elif "$inf" in raw:
            raw = -math.inf if raw["$inf"] == "-" else math.inf
    if kind == "float":
        if raw is None or raw == "":
            return math.nan
        try:
            return float(raw)
        except (TypeError, ValueError):
            return math.nan
    if kind == "int":
        try:
            return int(raw)
        except (TypeError, ValueError):
            return 0
    if kind == "bool":
        return raw is True or raw in ("true", "True")
    return "" if raw is None else str(raw)


class Schema:
    """Schema.fromJson (Schema.scala:140-170): component specs by name and by
    PyPSA list name."""

    def __init__(self, specs: List[ComponentSpec]):
        self.specs = specs
        self._by_name = {s.name: s for s in specs}
        self._by_list = {s.list_name: s for s in specs}

    def get(self, name: str) -> Optional[ComponentSpec]:
        return self._by_name.get(name) or self._by_list.get(name)

    def __getitem__(self, name: str) -> ComponentSpec:
        s = self.get(name)
        if s is None:
            raise KeyError(f"no component type '{name}'")
        return s

    @staticmethod
    def from_json(text: str) -> "Schema":
        root = json.loads(text)
        specs = []
        for cname in sorted(root):
            body = root[cname]
            attrs = {}
            for aname, a in body["attributes"].items():
                kind = _kind(a["type"])
                attrs[aname] = Attribute(aname, kind, _default(a.get("default"), kind),
                                         bool(a.get("varying", False)))
            specs.append(ComponentSpec(cname, body["list_name"], body.get("category") or "",
                                       attrs))
        return Schema(specs)

    @staticmethod
    def default() -> "Schema":
        """The schema of PyPSA 1.2.4 shipped with the package."""
        global _DEFAULT_SCHEMA
        if _DEFAULT_SCHEMA is None:
            with open(os.path.join(DATA_DIR, "schema.json")) as f:
                _DEFAULT_SCHEMA = Schema.from_json(f.read())
        return _DEFAULT_SCHEMA


_DEFAULT_SCHEMA: Optional[Schema] = None


# --------------------------------------------------------------------------
# tables


@dataclass
class Series:
    """TimeSeries (ComponentTable.scala:46-67): overrides for the entities
    that vary, values[snapshot, position]."""

    entities: List[str]
    values: np.ndarray

    def __post_init__(self):
        self._pos = {e: i for i, e in enumerate(self.entities)}

    def covers(self, entity: str) -> bool:
        return entity in self._pos

    def get(self, entity: str, t: int) -> Optional[float]:
        i = self._pos.get(entity)
        return None if i is None else float(self.values[t, i])


@dataclass
class Table:
    """ComponentTable (ComponentTable.scala:69-136)."""

    spec: ComponentSpec
    ids: List[str]
    static: Dict[str, list] = field(default_factory=dict)
    kinds: Dict[str, str] = field(default_factory=dict)
    series: Dict[str, Series] = field(default_factory=dict)

    def __post_init__(self):
        self._row = {e: i for i, e in enumerate(self.ids)}

    @property
    def name(self) -> str:
        return self.spec.name

    @property
    def size(self) -> int:
        return len(self.ids)

    def has(self, entity: str) -> bool:
        return entity in self._row

    def row(self, entity: str) -> int:
        try:
            return self._row[entity]
        except KeyError:
            raise KeyError(f"{self.spec.list_name} has no entity '{entity}'") from None

    def _column(self, attribute: str, entity: str, kind: str):
        have = self.kinds[attribute]
        if have != kind:
            raise ValueError(f"{attribute} is {have}, not {kind}")
        return self.static[attribute][self.row(entity)]

    def _spec_default(self, attribute: str) -> Attribute:
        a = self.spec.attribute(attribute)
        if a is None:
            raise KeyError(f"{self.spec.list_name} has no attribute '{attribute}'")
        return a

    def float(self, attribute: str, entity: str) -> float:
        """A present column must be float; an absent one reads the schema
        default as a number (ComponentTable.scala:83-90, :129-136)."""
        if attribute in self.static:
            return float(self._column(attribute, entity, "float"))
        d = self._spec_default(attribute).default
        if isinstance(d, bool) or isinstance(d, str):
            return math.nan
        return float(d)

    def string(self, attribute: str, entity: str) -> str:
        if attribute in self.static:
            return self._column(attribute, entity, "str")
        d = self._spec_default(attribute).default
        return d if isinstance(d, str) else ""

    def bool(self, attribute: str, entity: str) -> bool:
        if attribute in self.static:
            return bool(self._column(attribute, entity, "bool"))
        return self._spec_default(attribute).default is True

    def int(self, attribute: str, entity: str) -> int:
        if attribute in self.static:
            return int(self._column(attribute, entity, "int"))
        d = self._spec_default(attribute).default
        return d if isinstance(d, int) and not isinstance(d, bool) else 0

    def value_at(self, attribute: str, entity: str, t: int) -> float:
        """ComponentTable.valueAt (:117-127): override, static value, default."""
        s = self.series.get(attribute)
        if s is not None:
            v = s.get(entity, t)
            if v is not None:
                return v
        return self.float(attribute, entity)

    def with_floats(self, columns: Dict[str, list]) -> "Table":
        static = dict(self.static)
        kinds = dict(self.kinds)
        for k, v in columns.items():
            static[k] = list(v)
            kinds[k] = "float"
        return Table(self.spec, list(self.ids), static, kinds, dict(self.series))


@dataclass
class Network:
    """Network (ComponentTable.scala:145-199)."""

    name: str
    schema: Schema
    snapshots: List[str]
    tables: Dict[str, Table]
    weightings: Dict[str, np.ndarray]

    def weighting(self, kind: str, t: int) -> float:
        w = self.weightings.get(kind)
        return 1.0 if w is None else float(w[t])

    def table(self, component: str) -> Optional[Table]:
        t = self.tables.get(component)
        if t is None:
            spec = self.schema.get(component)
            if spec is not None:
                t = self.tables.get(spec.name)
        return t

    def require(self, component: str) -> Table:
        t = self.table(component)
        if t is None:
            raise KeyError(f"network '{self.name}' has no {component}")
        return t

    def with_table(self, t: Table) -> "Network":
        tables = dict(self.tables)
        tables[t.name] = t
        return Network(self.name, self.schema, self.snapshots, tables, self.weightings)


# --------------------------------------------------------------------------
# CSV reading


def _rows(path: str) -> List[List[str]]:
    """Quoted fields with doubled quotes, blank lines skipped
    (CsvReader.parse/splitLine, CsvReader.scala:258-294)."""
    if not os.path.exists(path):
        return []
    with open(path, newline="") as f:
        return [r for r in csv.reader(f) if r]


def _parse_float(cell: str, where: str) -> float:
    """CsvReader.parseFloat (:224-238): Python float syntax plus inf/nan words."""
    t = cell.strip()
    if not t:
        return math.nan
    low = t.lower()
    if low in ("inf", "+inf", "infinity", "+infinity"):
        return math.inf
    if low in ("-inf", "-infinity"):
        return -math.inf
    if low in ("nan", "na"):
        return math.nan
    try:
        return float(t)
    except ValueError:
        raise MalformedNetwork(f"{where}: '{cell}' is not a number") from None


def _typed_column(attr: Attribute, cells: List[str], where: str):
    """CsvReader.column (:144-181): a blank cell is the schema default."""
    d = attr.default
    if attr.kind == "float":
        return [d if not c.strip() else _parse_float(c, where) for c in cells]
    if attr.kind == "int":
        out = []
        for c in cells:
            t = c.strip()
            if not t:
                out.append(d)
                continue
            try:
                out.append(int(t))
            except ValueError:
                try:
                    out.append(int(float(t)))
                except ValueError:
                    out.append(d)
        return out
    if attr.kind == "bool":
        return [d if not c.strip() else c.strip().lower() in ("true", "1", "1.0") for c in cells]
    return [c.strip() or d for c in cells]


def _inferred_column(name: str, cells: List[str]):
    """CsvReader.inferColumn (:188-204): a custom column's kind from its
    values; port columns are always bus names."""
    t = [c.strip() for c in cells]
    present = [c for c in t if c]
    if is_port_column(name):
        return "str", t
    if present and all(c in ("True", "False") for c in present):
        return "bool", [c == "True" for c in t]
    if present:
        try:
            return "float", [math.nan if not c else float(c) for c in t]
        except ValueError:
            pass
    return "str", t


def _static_table(rows: List[List[str]], spec: ComponentSpec, where: str) -> Table:
    if not rows or not rows[0]:
        return Table(spec, [])
    header, body = rows[0], rows[1:]
    ids = [r[0] if r else "" for r in body]
    static, kinds = {}, {}
    for j, cname in enumerate(header):
        if j == 0:
            continue
        cells = [r[j] if j < len(r) else "" for r in body]
        a = spec.attribute(cname)
        if a is not None:
            static[cname] = _typed_column(a, cells, f"{where}.{cname}")
            kinds[cname] = a.kind
        else:
            kinds[cname], static[cname] = _inferred_column(cname, cells)
    return Table(spec, ids, static, kinds)


def read_table_csv(path: str, spec: ComponentSpec) -> Table:
    """One component table from a CSV file (CsvReader.table, :123-126)."""
    return _static_table(_rows(path), spec, spec.list_name)


def _snapshots(path: str):
    """CsvReader.readSnapshots (:94-117): `,snapshot,objective,stores,
    generators`; the label is the `snapshot` column, not the index."""
    rows = _rows(path)
    if not rows:
        return [], {}
    header, body = rows[0], rows[1:]
    label_at = header.index("snapshot") if "snapshot" in header else (1 if len(header) > 1 else 0)
    labels = [r[label_at] if label_at < len(r) else "" for r in body]
    weights = {}
    for i, name in enumerate(header):
        if i == label_at or not name:
            continue
        vals = []
        for r in body:
            try:
                vals.append(float(r[i].strip()) if i < len(r) else 1.0)
            except ValueError:
                vals.append(1.0)
        weights[name] = np.array(vals, dtype=np.float64)
    return labels, weights


def read_folder(directory: str, schema: Optional[Schema] = None, name: str = "") -> Network:
    """CsvReader.read (CsvReader.scala:34-79)."""
    if not os.path.isdir(directory):
        raise MalformedNetwork(f"{directory} is not a directory")
    schema = schema or Schema.default()
    files = sorted(f for f in os.listdir(directory) if f.endswith(".csv"))
    reserved = {"snapshots.csv", "network.csv", "investment_periods.csv"}
    snapshots, weights = _snapshots(os.path.join(directory, "snapshots.csv"))
    static_files = [f for f in files if "-" not in f and f not in reserved]
    series_files = [f for f in files if "-" in f and f not in reserved]
    tables: Dict[str, Table] = {}
    for f in static_files:
        list_name = f[:-4]
        spec = schema.get(list_name)
        if spec is None:
            continue
        table = read_table_csv(os.path.join(directory, f), spec)
        for sf in series_files:
            stem = sf[:-4]
            if stem[: stem.index("-")] != list_name:
                continue
            attribute = stem[stem.index("-") + 1:]
            rows = _rows(os.path.join(directory, sf))
            if not rows:
                continue
            entities = [e.strip() for e in rows[0][1:]]
            body = rows[1:]
            if len(body) != len(snapshots):
                raise MalformedNetwork(f"{list_name}-{attribute} has {len(body)} rows but the "
                                       f"network has {len(snapshots)} snapshots")
            vals = np.array([[_parse_float(r[i + 1] if i + 1 < len(r) else "", attribute)
                              for i in range(len(entities))] for r in body], dtype=np.float64)
            table.series[attribute] = Series(entities, vals.reshape(len(body), len(entities)))
        tables[spec.name] = table
    return Network(name or os.path.basename(os.path.normpath(directory)), schema, snapshots,
                   tables, weights)


# --------------------------------------------------------------------------
# topology


def role(spec: ComponentSpec) -> str:
    """Role.of (Topology.scala:41-48), from PyPSA's own category."""
    if spec.name == "Bus":
        return "bus"
    return {"passive_branch": "passive", "controllable_branch": "controllable",
            "controllable_one_port": "attached", "passive_one_port": "attached"}.get(
                spec.category, "none")


def branch_ports(table: Table) -> List[str]:
    """Topology.portsOf (Topology.scala:212-226): port columns in port order;
    both bus0 and bus1 are required."""
    ports = sorted((k for k in table.static if is_port_column(k)), key=lambda p: int(p[3:]))
    if table.size > 0:
        for required in ("bus0", "bus1"):
            if required not in ports:
                raise MalformedNetwork(f"{table.spec.list_name}: the '{required}' column is "
                                       "missing, and every branch joins two buses")
    return ports


def port_efficiency(table: Table, entity: str, port: str, t: int) -> float:
    """Topology.portEfficiency (Topology.scala:182-210): `efficiency` for
    bus1, `efficiency<i>` for later ports, 1 where unset or not finite."""
    attribute = "efficiency" if port == "bus1" else f"efficiency{port[3:]}"
    s = table.series.get(attribute)
    if s is not None and s.covers(entity):
        v = s.get(entity, t)
        return v if math.isfinite(v) else 1.0
    if table.spec.attribute(attribute) is not None or attribute in table.static:
        if attribute in table.static and table.kinds[attribute] != "float":
            return 1.0
        v = table.value_at(attribute, entity, t)
        return v if math.isfinite(v) else 1.0
    return 1.0


def dangling_bus_references(net: Network):
    """Topology.danglingBusReferences (Topology.scala:281-299): branch ports
    and one-port buses naming no bus. A blank extra port (bus2 and later) is
    PyPSA's default and is not dangling."""
    known = set(net.table("Bus").ids) if net.table("Bus") else set()
    out = []
    for table in net.tables.values():
        r = role(table.spec)
        if r in ("passive", "controllable"):
            for port in branch_ports(table):
                declared = table.spec.attribute(port) is not None
                for e in table.ids:
                    b = table.string(port, e)
                    if (b and b not in known) or (not b and declared):
                        out.append((table.name, e, port, b))
    for table in net.tables.values():
        if role(table.spec) == "attached" and table.spec.attribute("bus") is not None:
            for e in table.ids:
                b = table.string("bus", e)
                if b not in known:
                    out.append((table.name, e, "bus", b))
    return out


def sub_networks(net: Network):
    """Topology.subNetworks (Topology.scala:101-165): islands over passive
    branches, each labelled with the carrier of its smallest bus name.
    Returns [(carrier, sorted buses)] ordered by first bus."""
    buses = net.table("Bus")
    if buses is None:
        return []
    parent = {b: b for b in buses.ids}

    def find(x):
        root = x
        while parent[root] != root:
            root = parent[root]
        while parent[x] != root:
            parent[x], x = root, parent[x]
        return root

    for table in net.tables.values():
        if role(table.spec) != "passive":
            continue
        ports = branch_ports(table)
        for e in table.ids:
            ends = [table.string(p, e) for p in ports]
            ends = [b for b in ends if b]
            for b in ends:
                if b not in parent:
                    raise MalformedNetwork(f"{table.name} '{e}' names bus '{b}', which the network does not define")
            for i in range(len(ends)):
                for j in range(i + 1, len(ends)):
                    ra, rb = find(ends[i]), find(ends[j])
                    if ra != rb:
                        parent[ra] = rb
    islands: Dict[str, List[str]] = {}
    for b in buses.ids:
        islands.setdefault(find(b), []).append(b)
    out = []
    for members in islands.values():
        members = sorted(members)
        out.append((buses.string("carrier", members[0]), members))
    out.sort(key=lambda x: x[1][0])
    return out
