"""The helpers fp_filter imports from the reference's multi_filter.py (pycbc_parser :29-43, setopt :46-52,
getopt :55-56, dataset_argv :59-66, edges_per_template :264-282), without PyCBC: the parser holds the
pycbc_inspiral options this path reads, and edges_per_template runs on the GPU edges kernel."""
from __future__ import annotations

import argparse

import numpy as np

from chisq_edges import edges_blocked


def pycbc_parser():
    """The pycbc_inspiral options the filtering path reads (multi_filter.py:29-43); unknown options pass through
    parse_known_args as in the reference."""
    p = argparse.ArgumentParser(add_help=False)
    a = p.add_argument
    a('--bank-file'); a('--approximant', nargs='+')
    a('--order', type=int, default=-1)
    a('--low-frequency-cutoff', type=float)
    a('--snr-threshold', type=float)
    a('--cluster-window', type=float)
    a('--chisq-bins', default='0')
    a('--chisq-snr-threshold', type=float)
    a('--sample-rate', type=float)
    a('--segment-length', type=int); a('--segment-start-pad', type=int); a('--segment-end-pad', type=int)
    a('--psd-estimation'); a('--psd-segment-length', type=float); a('--psd-segment-stride', type=float)
    a('--psd-num-segments', type=int); a('--psd-inverse-length', type=float); a('--invpsd-trunc-method')
    a('--autogating-threshold', type=float); a('--autogating-cluster', type=float, default=5.0)
    a('--autogating-width', type=float, default=0.25); a('--autogating-taper', type=float, default=0.25)
    a('--autogating-pad', type=float, default=16)
    a('--newsnr-threshold', type=float); a('--keep-loudest-interval', type=float)
    a('--keep-loudest-num', type=int); a('--keep-loudest-stat')
    return p


def setopt(av, k, v):
    av = list(av)
    if k in av:
        av[av.index(k) + 1] = v
    else:
        av += [k, v]
    return av


def getopt(av, k, default):
    return float(av[av.index(k) + 1]) if k in av else default


def dataset_argv(base, ifo, s, e, store):
    pad = getopt(base, '--pad-data', 8); sp = getopt(base, '--segment-start-pad', 144)
    ep = getopt(base, '--segment-end-pad', 16)
    av = base
    for k, v in (('--hdf-store', store), ('--channel-name', f'{ifo}:GWOSC_STRAIN'),
                 ('--gps-start-time', str(int(s + pad))), ('--gps-end-time', str(int(e - pad))),
                 ('--trig-start-time', str(int(s + pad + sp))), ('--trig-end-time', str(int(e - pad - ep)))):
        av = setopt(av, k, v)
    return av


def edges_per_template(xp, rows, nb_, kmin, k1, Hd, ips):
    """Equal-power chi-squared bin edges for every peak (multi_filter.py:264-282): the GPU float64 edges kernel,
    once per distinct row. (npk, B + 1) int64."""
    return edges_blocked(xp, np.asarray(rows), np.asarray(nb_), kmin, np.asarray(k1), Hd, ips)
