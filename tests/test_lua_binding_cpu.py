"""CPU checks of the LuaJIT binding (hydro-cl-lua_amd/lua/hydro/): LuaJIT is
not in this image, so the Lua files are checked by what fixes their meaning:

- the header rewrite that hydro/hip/lib.lua hands to ffi.cdef (its gsub
  patterns, applied here with the same semantics) leaves plain C declarations
  that gcc accepts as C99, with no preprocessor line or C++ guard left;
- the name tables of lib.lua (limiters, integrators, fluxes, reconstructions,
  boundaries, equations) give the same indices as the header's enums, which
  the Python mirror (hydro._cfg) uses too;
- every C function the Lua files call is declared by the header.
"""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LUA = os.path.join(REPO, 'hydro-cl-lua_amd', 'lua', 'hydro')
HEADER = os.path.join(REPO, 'include', 'hydro_hip.h')


def _lua(name):
    return open(os.path.join(LUA, name)).read()


def _cdef_text(header):
    """lib.lua's cdefText: Lua's gsub replaces every non-overlapping match,
    left to right; '#[^\\n]*\\n', 'HYDRO_MAX_STATES', 'extern "C" {' and
    '\\n}%s*\\n' (%s: any whitespace, newlines included) read the same as
    these regular expressions"""
    s = re.sub(r'#[^\n]*\n', '\n', header)
    s = s.replace('HYDRO_MAX_STATES', '12')
    s = s.replace('extern "C" {', '')
    s = re.sub(r'\n\}\s*\n', '\n', s)
    return s


def test_lua_rewrite_patterns_are_the_ones_tested():
    src = _lua('hip/lib.lua')
    for pat in ("gsub('#[^\\n]*\\n', '\\n')", "gsub('HYDRO_MAX_STATES', '12')", """gsub('extern "C" {', '')""",
                "gsub('\\n}%s*\\n', '\\n')"):
        assert pat in src, pat


def test_cdef_text_is_plain_c99(tmp_path):
    text = _cdef_text(open(HEADER).read())
    assert not re.search(r'^\s*#', text, re.M)
    assert 'extern' not in text and '__cplusplus' not in text
    f = tmp_path / 'cdef.c'
    # ffi.cdef knows the stdint / stddef types; C needs their headers
    f.write_text('#include <stddef.h>\n#include <stdint.h>\n' + text)
    r = subprocess.run(['gcc', '-std=c99', '-pedantic-errors', '-Wall', '-fsyntax-only', str(f)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _lua_list(src, name):
    m = re.search(r'M\.%s = indexOf\{(.*?)\n\}' % name, src, re.S)
    assert m, name
    return re.findall(r"'([^']*)'", m.group(1))


def _lua_map(src, name):
    m = re.search(r'M\.%s = \{(.*?)\}\n' % name, src, re.S)
    assert m, name
    out = {}
    for k1, k2, v in re.findall(r"(?:\['([^']*)'\]|(\w+)) = (\d+)", m.group(1)):
        out[k1 or k2] = int(v)
    return out


def _enum(header, prefix):
    """name -> value of the header's enum constants with this prefix"""
    vals, nxt = {}, 0
    for body in re.findall(r'enum\s*\{(.*?)\}', header, re.S):
        nxt = 0
        for item in re.sub(r'/\*.*?\*/', '', body, flags=re.S).split(','):
            item = item.strip()
            if not item:
                continue
            if '=' in item:
                k, v = [x.strip() for x in item.split('=')]
                nxt = int(v)
            else:
                k = item
            if k.startswith(prefix):
                vals[k] = nxt
            nxt += 1
    return vals


def test_name_tables_match_header_and_python_mirror():
    from hydro import _cfg
    src = _lua('hip/lib.lua')
    header = open(HEADER).read()
    assert _lua_list(src, 'limiters') == _cfg.LIMITERS
    assert len(_cfg.LIMITERS) == _enum(header, 'HYDRO_NUM_LIMITERS')['HYDRO_NUM_LIMITERS']
    assert _lua_list(src, 'integrators') == _cfg.INTEGRATORS
    assert len(_cfg.INTEGRATORS) == _enum(header, 'HYDRO_NUM_INTEGRATORS')['HYDRO_NUM_INTEGRATORS']
    fluxes = _lua_map(src, 'fluxes')
    assert fluxes == {name: i for i, name in enumerate(_cfg.FLUXES)}
    hf = _enum(header, 'HYDRO_FLUX_')
    assert [hf[k] for k in ('HYDRO_FLUX_ROE', 'HYDRO_FLUX_HLL', 'HYDRO_FLUX_RUSANOV', 'HYDRO_FLUX_HLLC',
                            'HYDRO_FLUX_MARQUINA')] == [0, 1, 2, 3, 4]
    plms = _lua_map(src, 'plms')
    for i, name in enumerate(_cfg.PLMS):
        if name is not None:
            assert plms[name] == i
    assert plms['piecewise constant'] == 0
    assert _enum(header, 'HYDRO_PLM_')['HYDRO_PLM_EIG_PRIM_REF'] == len(_cfg.PLMS) - 1
    assert _lua_map(src, 'boundaries') == {name: i for i, name in enumerate(_cfg.BOUNDARIES)}
    eq = _lua_map(src, 'equations')
    he = _enum(header, 'HYDRO_EQN_')
    assert eq == {'euler': he['HYDRO_EQN_EULER'], 'glm-maxwell': he['HYDRO_EQN_GLM_MAXWELL'],
                  'shallow-water': he['HYDRO_EQN_SHALLOW_WATER']}
    assert [_cfg.EQNS[i].replace('_', '-') for i in sorted(eq.values())] == sorted(eq, key=eq.get)


@pytest.mark.parametrize('name', ['hip/lib.lua', 'solver/hipfvsolver.lua'])
def test_lua_calls_only_declared_functions(name):
    header = open(HEADER).read()
    declared = set(re.findall(r'\b(hydro_\w+)\s*\(', header))
    called = set(re.findall(r'\bC\.(hydro_\w+)', _lua(name)))
    assert called, name
    assert called <= declared, called - declared


def test_solver_cons_t_sizes_match_header():
    """hipfvsolver.lua's numStates / numIntStates per equation are the
    library's (hydro_eqn_num_states / _int_states through the Python mirror)"""
    from hydro._cfg import EQN_NUM_STATES, EQN_NUM_INT_STATES
    src = _lua('solver/hipfvsolver.lua')
    found = re.findall(r"\[?'?([\w-]+)'?\]? = \{numStates = (\d+), numIntStates = (\d+)\}", src)
    assert len(found) == 3, found
    for name, S, q in found:
        key = name.replace('-', '_')
        assert EQN_NUM_STATES[key] == int(S) and EQN_NUM_INT_STATES[key] == int(q), name
