"""Run provenance for peak files and the analysis-history log (the reference's provenance.py:21-81): md5 of every
input file, the command line, md5 of each loaded module of this package, and one config_md5 over them, so that two
runs with equal config_md5 used identical inputs and code."""
from __future__ import annotations

import datetime
import hashlib
import json
import os
import socket
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PLACE_OPTS = {'--out', '--output', '--prep-cache', '--hdf-store', '--history'}


def md5(path, bs=1 << 22):
    if not path or not os.path.isfile(path):
        return None
    h = hashlib.md5()
    with open(path, 'rb') as f:
        for b in iter(lambda: f.read(bs), b''):
            h.update(b)
    return h.hexdigest()


def code_md5s():
    """md5 of every loaded module whose file lies in this package, and of the running script."""
    out = {}
    for m in list(sys.modules.values()):
        f = getattr(m, '__file__', None)
        if f and os.path.abspath(f).startswith(HERE + os.sep) and f.endswith('.py'):
            out[os.path.relpath(os.path.abspath(f), HERE)] = md5(f)
    main = os.path.abspath(sys.argv[0]) if sys.argv and sys.argv[0] else ''
    if main.endswith('.py'):
        out[os.path.basename(main)] = md5(main)
    return dict(sorted(out.items()))


def record(inputs, argv=None, dq_files=(), dq_flags=()):
    """inputs: {role: path}; dq_files: segment files that define the analysed time; dq_flags: gating options. The
    config_md5 ignores where things live (paths, output and cache directories), only what they contain."""
    p = dict(argv=list(sys.argv if argv is None else argv),
             inputs={k: dict(path=v, md5=md5(v)) for k, v in inputs.items() if v},
             code=code_md5s(),
             dq=dict(source='science segments of the datasets file + autogating',
                     files={os.path.basename(f): md5(f) for f in dq_files if f}, flags=list(dq_flags)))
    p['dq']['version'] = hashlib.md5(json.dumps(p['dq'], sort_keys=True).encode()).hexdigest()[:12]
    av = p['argv'][1:]
    av = [os.path.basename(x) if '/' in x else x for i, x in enumerate(av) if not (i and av[i - 1] in PLACE_OPTS)]
    key = dict(argv=av, inputs={k: v['md5'] for k, v in p['inputs'].items()}, code=p['code'], dq=p['dq']['version'])
    p['config_md5'] = hashlib.md5(json.dumps(key, sort_keys=True).encode()).hexdigest()
    return p


def log(history, stage, status, out=None, prov=None, **extra):
    """Append one JSON line to the analysis-history log (never rewritten)."""
    if not history:
        return
    e = dict(time=datetime.datetime.now(datetime.timezone.utc).isoformat(timespec='seconds'),
             host=socket.gethostname(), stage=stage, status=status, out=out,
             config_md5=(prov or {}).get('config_md5'), dq_version=((prov or {}).get('dq') or {}).get('version'),
             **extra)
    os.makedirs(os.path.dirname(os.path.abspath(history)), exist_ok=True)
    with open(history, 'a') as f:
        f.write(json.dumps(e, default=str) + '\n')
