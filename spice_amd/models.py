"""Model registry: ids of include/spice_hip.h and attribute names/types per model."""
import numpy as np

# name: (id, neuron attributes, synapse attributes); attribute = (name, numpy dtype)
MODELS = {
    "vogels_abbott": (0, [("V", np.float32), ("Gex", np.float32), ("Gin", np.float32), ("Twait", np.int32)], []),
    "brunel": (1, [("V", np.float32), ("Twait", np.int32)], []),
    "brunel_with_plasticity": (2, [("V", np.float32), ("Twait", np.int32)],
                               [("W", np.float32), ("Zpre", np.float32), ("Zpost", np.float32)]),
    "synth": (3, [("N", np.int32)], []),
    "brunel_3pop": (4, [("V", np.float32), ("Twait", np.int32)], []),
    "brunel_3pop_stdp": (5, [("V", np.float32), ("Twait", np.int32)], [("W", np.float32)]),
    # brunel_3pop_stdp without its refractory promise: the two-kernel step with deferred events
    "brunel_3pop_stdp_deferred": (6, [("V", np.float32), ("Twait", np.int32)], [("W", np.float32)]),
}

PLASTIC = {name for name, (_, _, syn) in MODELS.items() if syn}


def model_id(name):
    if name not in MODELS:
        raise ValueError(f"unknown model {name!r}; known: {sorted(MODELS)}")
    return MODELS[name][0]


def neuron_attrs(name):
    return MODELS[name][1]


def synapse_attrs(name):
    return MODELS[name][2]
