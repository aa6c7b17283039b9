"""Systems whose vdW pairs use every potential form of vdw_terms.f:189-396
(keys 1-9), for the tests of the non-Lennard-Jones forms.

Each Lennard-Jones pair (eps, sigma) of a synthetic system is given another
form with a well of about the same depth and place, so the chains stay
physical: 12-6 and n-m with the same well, a Buckingham and a
Born-Huggins-Meyer pair dominated by their exponential repulsion (their
dispersion terms are kept tiny, so the r -> 0 collapse of exp-6 forms lies
far inside any distance a trial reaches), a 12-10 hydrogen bond, the shifted
n-m form, Morse and WCA.  Pairs with a zero eps or sigma stay the
Lennard-Jones zero pair."""
import copy
import math

RM = 2.0 ** (1.0 / 6.0)


def key_params(key, eps, sig):
    """Parameters (prmvdw order, vdw_terms.f:225-392) of form `key` with a
    well of depth about eps at about 2^(1/6) sigma."""
    rm = RM * sig
    if key == 1:
        return [4.0 * eps * sig ** 12, 4.0 * eps * sig ** 6, 0.0, 0.0, 0.0]
    if key == 2:
        return [eps, sig, 0.0, 0.0, 0.0]
    if key == 3:
        return [eps, 10.0, 5.0, rm, 0.0]
    if key == 4:
        return [eps * 6.0 / 7.0 * math.exp(13.0), rm / 13.0, 1e-6 * eps * rm ** 6, 0.0, 0.0]
    if key == 5:
        return [eps * 6.0 / 7.0, 13.0 / rm, rm, 1e-6 * eps * rm ** 6, 1e-6 * eps * rm ** 8]
    if key == 6:
        return [5.0 * eps * rm ** 12, 6.0 * eps * rm ** 10, 0.0, 0.0, 0.0]
    if key == 7:
        return [eps, 12.0, 6.0, rm, 0.0]
    if key == 8:
        return [eps, rm, 6.0 / rm, 0.0, 0.0]
    if key == 9:
        return [eps, sig, 0.0, 0.0, 0.0]
    raise ValueError(key)


def keyed_arrays(A, keys=tuple(range(1, 10)), first=0):
    """A copy of the system arrays A whose vdW pairs take the forms `keys`
    in turn (the pair of index p gets keys[(p + first) % len(keys)]).
    Returns (arrays, the keys used)."""
    B = copy.deepcopy(A)
    out, used = [], []
    p = 0
    for (ti, tj, key, prm) in A["vdw"]:
        eps, sig = prm[0], prm[1]
        if key != 2 or eps == 0.0 or sig == 0.0:
            out.append((ti, tj, key, list(prm)))
            continue
        k = keys[(p + first) % len(keys)]
        p += 1
        out.append((ti, tj, k, key_params(k, eps, sig)))
        used.append(k)
    B["vdw"] = out
    return B, sorted(set(used))
