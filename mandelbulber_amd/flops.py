"""Algorithmic work of a render, the numerator of the roofline figure.

Counting rule: +, -, *, /, sqrt = 1 flop each, FMA = 2, each transcendental
(sin, cos, asin, atan2, pow, log) = 1.  Per formula iteration of
Compute<normal> (src/fractal.cpp:211-225, :369-529):
  Mandelbulb (trig_optim)  28  (21 arithmetic + 7 transcendental)
  Mandelbox (tglad)        36
  Menger sponge            19
per DE call tail (escape test and distance): Mandelbulb 4, Mandelbox 2,
Menger 2; per march or shading step 8 (threshold, point = p + v*t, step).
The counts of iterations and DE calls are measured by the kernels.
"""

F_ITER = {2: 28, 8: 36, 7: 19}
F_TAIL = {2: 4, 8: 2, 7: 2}
F_STEP = 8


def algorithmic_flops(scene, stats, kernel):
    f_iter = F_ITER[scene.formula]
    f_call = F_TAIL[scene.formula] + F_STEP
    if kernel == "ao":
        return stats.ao_iterations * f_iter + stats.ao_de_calls * f_call
    if kernel == "shade":
        return stats.shade_iterations * f_iter + stats.shade_de_calls * f_call
    if kernel == "march":
        return stats.march_iterations * f_iter + stats.march_de_calls * f_call
    return (stats.march_iterations + stats.shade_iterations) * f_iter + \
        (stats.march_de_calls + stats.shade_de_calls) * f_call
