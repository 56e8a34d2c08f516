"""Drop-in for the reference's Python `pape` extension module
(trunk/python_extension/pape_module.c:301-473): a graph held by the module,
costs replaced in place, and the predecessor-node list of a d'Esopo-Pape
search -- here the engine's all-origin search run for one origin.

    build_adjlist(graph_dict, num_nodes, first_thru_node)
    update_weights(list_of_costs)       # edge order: sorted by (from, to)
    pape(start_node) -> list[num_nodes] of predecessor nodes (-1: none)

Errors as in the reference: IndexError for a weight list of the wrong
length or a start node >= num_nodes.  Node ids are 1-based (node 0 has no
links), as in the TNTP data the reference feeds it.
"""
from __future__ import annotations

import numpy as np

from .. import synth
from ..engine import TapEngine

_state = {"engine": None, "num_nodes": 0, "tail": None, "num_edges": 0}


def build_adjlist(graph_dict, num_nodes, first_thru_node):
    """graph_dict[u][v] = cost; links sorted by (u, v) as the reference sorts
    them (pape_module.c:196-283)."""
    edges = sorted((int(u), int(v), float(c)) for u, nb in graph_dict.items()
                   for v, c in nb.items())
    if edges and min(min(u, v) for u, v, _ in edges) < 1:
        raise ValueError("node ids must be >= 1")
    if edges and max(max(u, v) for u, v, _ in edges) >= num_nodes:
        raise IndexError("node id >= num_nodes")
    E = len(edges)
    init = np.array([e[0] for e in edges], dtype=np.int64)
    term = np.array([e[1] for e in edges], dtype=np.int64)
    cost = np.array([e[2] for e in edges], dtype=np.float64)
    net = synth.Network(num_zones=0, num_nodes=int(num_nodes) - 1,
                        first_thru_node=int(first_thru_node), init=init, term=term,
                        capacity=np.ones(E), length=np.ones(E), fftime=cost, B=np.zeros(E),
                        power=np.ones(E), speed=np.zeros(E), toll=np.zeros(E),
                        ltype=np.ones(E, np.int16))
    dem = synth.Demand(0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0))
    if _state["engine"] is not None:
        _state["engine"].close()
    _state.update(engine=TapEngine(net, dem), num_nodes=int(num_nodes), tail=init,
                  num_edges=E)


def update_weights(weights):
    if len(weights) != _state["num_edges"]:
        raise IndexError("list length != num edges")
    _state["engine"].set_costs(np.asarray(weights, dtype=np.float64))


def pape(start_node):
    if start_node >= _state["num_nodes"]:
        raise IndexError("start_node >= num_nodes")
    _, p = _state["engine"].sssp([int(start_node)])
    links = p[:, 0]
    out = np.full(_state["num_nodes"], -1, dtype=np.int64)
    has = links >= 0
    out[has] = _state["tail"][links[has]]
    return out.tolist()
