"""JSON -> attribute config (reference ``src/utility.py:116-172`` get_config)."""
import json
import types


def _obj(d):
    if isinstance(d, dict):
        return types.SimpleNamespace(**{k: _obj(v) for k, v in d.items()})
    return d


def get_config(path_or_dict):
    if isinstance(path_or_dict, dict):
        return _obj(path_or_dict)
    with open(path_or_dict) as fh:
        return _obj(json.load(fh))
