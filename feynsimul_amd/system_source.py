"""Generation of the system section compiled into the sampler module.

The reference pastes the system's strings into its OpenCL template as
``potential(DOF_ARGUMENT_DECL)`` returning the potential expression,
``operator(...)`` doing ``opAccum[i] += <op_i>`` and ``correlator(...)`` doing
``corProd[i] *= <corr_i>`` (FeynSimul/kernel.py:253-285, kernel.c:108-145).
This module emits the same three functions (``operator`` is a C++ keyword,
so it is ``fs_operator`` here) plus array-argument wrappers that the samplers
call with the coordinates of one slice held in registers.  ``fs_hist_obs``
evaluates the scalar observables that the wave sampler bins into one
histogram (a pair-distance histogram, for example).
"""


def system_section(system, operators=(), correlators=(), observables=()):
    dof = int(system.DOF)
    decl = ", ".join("FLOAT_TYPE x%d" % (i + 1) for i in range(dof))
    call = ", ".join("x[%d]" % i for i in range(dof))
    out = []
    out.append(getattr(system, "userCode", "") or "")
    out.append("inline FLOAT_TYPE potential(%s)\n{\n    return %s;\n}" % (decl, system.potential))
    body = "".join("    opAccum[%d] += %s;\n" % (i, op) for i, op in enumerate(operators))
    out.append("inline void fs_operator(%s, FLOAT_TYPE *opAccum)\n{\n%s}" % (decl, body))
    body = "".join("    corProd[%d] *= %s;\n" % (i, c) for i, c in enumerate(correlators))
    out.append("inline void fs_correlator(%s, FLOAT_TYPE *corProd)\n{\n%s}" % (decl, body))
    body = "".join("    obs[%d] = %s;\n" % (i, o) for i, o in enumerate(observables))
    out.append("inline void fs_hist_obs(%s, FLOAT_TYPE *obs)\n{\n%s}" % (decl, body))
    out.append("inline FLOAT_TYPE fs_potential_v(const FLOAT_TYPE *x) { return potential(%s); }" % call)
    out.append("inline void fs_hist_obs_v(const FLOAT_TYPE *x, FLOAT_TYPE *o) { fs_hist_obs(%s, o); }" % call)
    out.append("inline void fs_operator_v(const FLOAT_TYPE *x, FLOAT_TYPE *a) { fs_operator(%s, a); }" % call)
    out.append("inline void fs_correlator_v(const FLOAT_TYPE *x, FLOAT_TYPE *p) { fs_correlator(%s, p); }" % call)
    return "\n".join(out) + "\n"
