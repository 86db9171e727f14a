"""Host-side summary of the learned-vs-plain evaluation: each method's time to
the plain run's error level, failures counted as infinite ratios in the
medians, and the ratio against the plain run with its own stop kept apart."""
import math

from airtoolsii_amd import learned as Lm


def _row(t_l, t_p, k_l, k_p, plain_s, kg=None):
    return {"plain_err": 0.1, "learned_err": 0.09, "plain_s": plain_s, "learned_s": 1.0,
            "time_to_plain_error": t_l, "plain_time_to_err": t_p, "iters_to_plain_err": k_l,
            "plain_iters_to_err": k_p, "plain_iters": k_p + 2, "learned_iters": 12, "plain_setup_s": 0.1,
            "learned_setup_s": 0.1, "greedy_iters_to_plain_err": kg}


def test_summary_medians_count_failures():
    rows = [_row(1.0, 2.0, 10, 20, 2.5, 8), _row(None, 2.0, None, 20, 2.5, 9), _row(1.2, 2.0, 12, 20, 2.5, 7)]
    s = Lm.summarize(rows)
    # time ratios 0.5, inf, 0.6 -> median 0.6; iterations 0.5, inf, 0.6
    assert math.isclose(s["time_to_plain_error_over_plain_median"], 0.6)
    assert math.isclose(s["iters_to_plain_err_over_plain_median"], 0.6)
    assert math.isclose(s["reached_plain_err_frac"], 2 / 3)
    # over the problems that reach the error: 0.5 and 0.6
    assert math.isclose(s["time_to_plain_error_over_plain_median_reached"], 0.55)
    # against the plain run with its own stop (2.5 s): 0.4, inf, 0.48
    assert math.isclose(s["time_to_plain_error_over_plain_dp_run_median"], 0.48)
    # greedy iterations over the plain iterations to the error: 0.4, 0.45, 0.35
    assert math.isclose(s["greedy_iters_to_plain_err_over_plain_median"], 0.4)


def test_summary_median_on_a_failure_is_null():
    rows = [_row(None, 2.0, None, 20, 2.5), _row(None, 2.0, None, 20, 2.5), _row(1.0, 2.0, 10, 20, 2.5)]
    s = Lm.summarize(rows)
    assert s["time_to_plain_error_over_plain_median"] is None
    assert s["reached_plain_err_frac"] == 1 / 3
