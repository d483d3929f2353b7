"""Host-side v4 optimizers that drive the device metric.

The optimizer state (a few vectors of P doubles) lives on the host; each
iteration is one device GetValueAndDerivative plus O(P) host arithmetic,
exactly the reference's split between optimizer and metric.

* ``RegularStepGradientDescentOptimizerv4`` follows
  Modules/Numerics/Optimizersv4/include/itkRegularStepGradientDescentOptimizerv4.hxx:25-171
  and the loop of itkGradientDescentOptimizerv4.hxx:72-165, including its
  handling of the previous gradient (re-multiplied by 1/scales every step).
* ``LBFGSBOptimizerv4`` runs the L-BFGS-B algorithm (Byrd, Lu, Nocedal & Zhu
  1995) through SciPy's implementation of the same published code the
  reference vendors (Modules/ThirdParty/VNL ... lbfgsb.c, wrapped by
  itkLBFGSBOptimizerv4.cxx).
"""
from __future__ import annotations

import ctypes
import math
from enum import Enum

import numpy as np


class StopCondition(Enum):
    MAXIMUM_NUMBER_OF_ITERATIONS = "MAXIMUM_NUMBER_OF_ITERATIONS"
    GRADIENT_MAGNITUDE_TOLERANCE = "GRADIENT_MAGNITUDE_TOLERANCE"
    STEP_TOO_SMALL = "STEP_TOO_SMALL"
    COSTFUNCTION_ERROR = "COSTFUNCTION_ERROR"
    CONVERGENCE_CHECKER_PASSED = "CONVERGENCE_CHECKER_PASSED"
    OTHER_ERROR = "OTHER_ERROR"


def compensated_sum(values) -> float:
    """CompensatedSummation (Modules/Core/Common/include/itkCompensatedSummation.hxx:38-48)."""
    s = 0.0
    c = 0.0
    for x in values:
        y = float(x) - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


class _OptimizerBase:
    """ObjectToObjectOptimizerBaseTemplate's scales handling
    (Modules/Numerics/Optimizersv4/src/itkObjectToObjectOptimizerBase.cxx:80-160)."""

    def __init__(self, number_of_iterations: int, scales=None, scales_estimator=None):
        self.number_of_iterations = int(number_of_iterations)
        self.scales = None if scales is None else np.asarray(scales, dtype=np.float64)
        self.scales_estimator = scales_estimator
        self.do_estimate_scales = True
        self.metric = None
        self.current_iteration = 0
        self.current_value = float(np.finfo(np.float64).max)
        self.stop_condition = None
        self.stop_description = ""
        self.observers = []

    def SetMetric(self, metric):
        self.metric = metric

    def SetScales(self, scales):
        self.scales = np.asarray(scales, dtype=np.float64)

    def GetScales(self):
        return self.scales

    def SetScalesEstimator(self, estimator):
        self.scales_estimator = estimator

    def SetNumberOfIterations(self, n: int):
        self.number_of_iterations = int(n)

    def GetCurrentIteration(self) -> int:
        return self.current_iteration

    def GetCurrentMetricValue(self) -> float:
        return self.current_value

    def GetValue(self) -> float:
        return self.current_value

    def GetStopCondition(self):
        return self.stop_condition

    def GetStopConditionDescription(self) -> str:
        return self.stop_description

    def AddObserver(self, fn):
        """fn(optimizer) is called once per iteration (the reference's IterationEvent)."""
        self.observers.append(fn)

    def _prepare_scales(self):
        if self.metric is None:
            raise ValueError("metric is not set")
        n_local = self.metric.GetNumberOfLocalParameters()
        if self.do_estimate_scales and self.scales_estimator is not None:
            self.scales = np.asarray(self.scales_estimator.EstimateScales(), dtype=np.float64)
        if self.scales is not None:
            if self.scales.size != n_local:
                raise ValueError("size of scales (%d) must equal number of local parameters (%d)"
                                 % (self.scales.size, n_local))
            if np.any(self.scales <= np.finfo(np.float64).eps):
                raise ValueError("m_Scales values must be > epsilon.")
            self.scales_are_identity = bool(np.all(np.abs(1.0 - self.scales) <= 0.01))
        else:
            self.scales = np.ones(n_local)
            self.scales_are_identity = True


class RegularStepGradientDescentOptimizerv4(_OptimizerBase):
    def __init__(self, learning_rate: float = 1.0, minimum_step_length: float = 1e-4,
                 relaxation_factor: float = 0.5, gradient_magnitude_tolerance: float = 1e-4,
                 number_of_iterations: int = 100, scales=None, scales_estimator=None,
                 return_best_parameters_and_value: bool = False):
        super().__init__(number_of_iterations, scales, scales_estimator)
        self.learning_rate = float(learning_rate)
        self.minimum_step_length = float(minimum_step_length)
        self.relaxation_factor = float(relaxation_factor)
        self.gradient_magnitude_tolerance = float(gradient_magnitude_tolerance)
        self.return_best = bool(return_best_parameters_and_value)
        self.current_relaxation = 0.0
        self.gradient = None
        self.previous_gradient = None
        # The iterations run in the C ABI's batched loop (bit-identical to the
        # loop below) when the metric is the device metric of a transform of at
        # most 12 parameters and nothing needs a per-iteration callback.
        # native_runner(metric, options, state) may hand them to a serving lane.
        self.use_native_loop = True
        self.native_runner = None

    # itk setters
    def SetLearningRate(self, v: float):
        self.learning_rate = float(v)

    def SetMinimumStepLength(self, v: float):
        self.minimum_step_length = float(v)

    def SetRelaxationFactor(self, v: float):
        self.relaxation_factor = float(v)

    def SetGradientMagnitudeTolerance(self, v: float):
        self.gradient_magnitude_tolerance = float(v)

    def GetCurrentStepLength(self) -> float:
        return self.current_relaxation * self.learning_rate

    def StartOptimization(self, do_only_initialization: bool = False):
        self._prepare_scales()
        if self.gradient_magnitude_tolerance < 0.0:
            raise ValueError("Gradient magnitude tolerance must be greater or equal 0.0. Current value is %g"
                             % self.gradient_magnitude_tolerance)
        P = self.metric.GetNumberOfParameters()
        self.gradient = np.zeros(P)
        self.previous_gradient = np.zeros(P)
        self.current_relaxation = 1.0
        self.current_iteration = 0
        if self.return_best:
            self.best_parameters = self.metric.GetParameters()
            self.best_value = float(np.finfo(np.float64).max)
        if do_only_initialization:
            return
        if self._native_ok():
            self._run_native()
        else:
            self.ResumeOptimization()

    def _native_ok(self) -> bool:
        if not self.use_native_loop or self.observers or self.return_best:
            return False
        if self.metric.GetNumberOfParameters() > 12:
            return False
        if self.native_runner is not None:
            return True
        from .metric import MattesMutualInformationImageToImageMetricv4
        return type(self.metric) is MattesMutualInformationImageToImageMetricv4

    def _run_native(self):
        from . import _native as N
        from .batch import shared_evaluator

        opts = N.RegularStepOptions(self.learning_rate, self.minimum_step_length, self.relaxation_factor,
                                    self.gradient_magnitude_tolerance, self.number_of_iterations, 0)
        st = N.RegularStepState()
        sc = np.ascontiguousarray(self.scales, dtype=np.float64)
        L = N.load()
        N.check(L.mmi_regular_step_start(ctypes.byref(st), sc.ctypes.data_as(N._dp), sc.size))
        if self.native_runner is not None:
            st = self.native_runner(self.metric, opts, st)
        else:
            states = [st]
            shared_evaluator().regular_step([self.metric], opts, states, self.number_of_iterations + 1)
            st = states[0]
        P = self.metric.GetNumberOfParameters()
        self.current_iteration = int(st.iteration)
        self.current_value = float(st.value)
        self.gradient = np.array(st.gradient[:P])
        self.previous_gradient = np.array(st.previous_gradient[:P])
        self.current_relaxation = float(st.relaxation)
        self.stop_description = "RegularStepGradientDescentOptimizerv4: "
        if st.stop == N.MMI_STOP_METRIC_ERROR:
            self.stop_condition = StopCondition.COSTFUNCTION_ERROR
            self.stop_description += "Metric error during optimization"
            # the message the metric's own evaluation gave (its exception text)
            msg = N.load().mmi_context_last_error(self.metric._ctx).decode()
            raise N.MetricError(int(st.status), msg or "metric evaluation failed during optimization (status %d)"
                                % st.status)
        if st.stop == N.MMI_STOP_MAXIMUM_NUMBER_OF_ITERATIONS:
            self.stop_condition = StopCondition.MAXIMUM_NUMBER_OF_ITERATIONS
            self.stop_description += "Maximum number of iterations (%d) exceeded." % self.number_of_iterations
        elif st.stop == N.MMI_STOP_GRADIENT_MAGNITUDE_TOLERANCE:
            self.stop_condition = StopCondition.GRADIENT_MAGNITUDE_TOLERANCE
            self.stop_description += ("Gradient magnitude tolerance met after %d iterations. Gradient magnitude (%g) "
                                      "is less than gradient magnitude tolerance (%g)."
                                      % (st.iteration - 1, st.gradient_magnitude, self.gradient_magnitude_tolerance))
        elif st.stop == N.MMI_STOP_STEP_TOO_SMALL:
            self.stop_condition = StopCondition.STEP_TOO_SMALL
            self.stop_description += ("Step too small after %d iterations. Current step (%g) is less than minimum "
                                      "step (%g)." % (st.iteration - 1, st.relaxation * self.learning_rate,
                                                      self.minimum_step_length))

    def StopOptimization(self):
        if self.return_best:
            self.metric.SetParameters(self.best_parameters)
            self.current_value = self.best_value
        self._stop = True

    def ResumeOptimization(self):
        self.stop_description = "RegularStepGradientDescentOptimizerv4: "
        self._stop = False
        while not self._stop:
            if self.current_iteration >= self.number_of_iterations:
                self.stop_description += "Maximum number of iterations (%d) exceeded." % self.number_of_iterations
                self.stop_condition = StopCondition.MAXIMUM_NUMBER_OF_ITERATIONS
                self.StopOptimization()
                break
            self.previous_gradient, self.gradient = self.gradient, self.previous_gradient
            try:
                self.current_value, g = self.metric.GetValueAndDerivative()
            except Exception:
                self.stop_condition = StopCondition.COSTFUNCTION_ERROR
                self.stop_description += "Metric error during optimization"
                self.StopOptimization()
                raise
            self.gradient = np.array(g, dtype=np.float64)
            for fn in self.observers:
                fn(self)
            if self._stop:
                break
            self.AdvanceOneStep()
            if self.return_best and self.current_value < self.best_value:
                self.best_value = self.current_value
                self.best_parameters = self.metric.GetParameters()
            self.current_iteration += 1

    def _modify_gradient_by_scales(self):
        if self.scales_are_identity or self.gradient.size == 0:
            return
        factor = 1.0 / self.scales
        idx = np.arange(self.gradient.size) % self.scales.size
        self.gradient = self.gradient * factor[idx]
        self.previous_gradient = self.previous_gradient * factor[idx]

    def AdvanceOneStep(self):
        if self.scales.size != self.gradient.size and not self.metric.HasLocalSupport():
            raise ValueError("The size of Scales is %d, but the NumberOfParameters for the CostFunction is %d."
                             % (self.scales.size, self.gradient.size))
        if self.relaxation_factor < 0.0:
            raise ValueError("Relaxation factor must be positive. Current value is %g" % self.relaxation_factor)
        if self.relaxation_factor >= 1.0:
            raise ValueError("Relaxation factor must be less than 1.0. Current value is %g" % self.relaxation_factor)
        self._modify_gradient_by_scales()
        magnitude = math.sqrt(compensated_sum(self.gradient * self.gradient))
        if magnitude < self.gradient_magnitude_tolerance:
            self.stop_condition = StopCondition.GRADIENT_MAGNITUDE_TOLERANCE
            self.stop_description += ("Gradient magnitude tolerance met after %d iterations. Gradient magnitude (%g) "
                                      "is less than gradient magnitude tolerance (%g)."
                                      % (self.current_iteration, magnitude, self.gradient_magnitude_tolerance))
            self.StopOptimization()
            return
        scalar = compensated_sum(self.gradient * self.previous_gradient)
        if scalar < 0:
            self.current_relaxation *= self.relaxation_factor
        step = self.current_relaxation * self.learning_rate
        if step < self.minimum_step_length:
            self.stop_condition = StopCondition.STEP_TOO_SMALL
            self.stop_description += ("Step too small after %d iterations. Current step (%g) is less than minimum "
                                      "step (%g)." % (self.current_iteration, step, self.minimum_step_length))
            self.StopOptimization()
            return
        self.gradient = self.gradient * self.current_relaxation * self.learning_rate
        self.metric.UpdateTransformParameters(self.gradient, 1.0 / magnitude)


class LBFGSBOptimizerv4(_OptimizerBase):
    """Bounded limited-memory BFGS.  The metric value is minimised; the
    metric's derivative is -d(value)/dp (the v4 convention), so the gradient
    handed to L-BFGS-B is its negation.  Defaults follow itkLBFGSBOptimizerv4.h:183
    and itkLBFGSOptimizerBasev4.h:160-163 (CostFunctionConvergenceFactor 1e7,
    GradientConvergenceTolerance 1e-5, MaximumNumberOfCorrections 5,
    MaximumNumberOfFunctionEvaluations 2000)."""

    def __init__(self, number_of_iterations: int = 500, cost_function_convergence_factor: float = 1e7,
                 projected_gradient_tolerance: float = 1e-5, maximum_number_of_corrections: int = 5,
                 maximum_number_of_function_evaluations: int = 2000, lower_bound=None, upper_bound=None):
        super().__init__(number_of_iterations)
        self.factr = float(cost_function_convergence_factor)
        self.pgtol = float(projected_gradient_tolerance)
        self.m = int(maximum_number_of_corrections)
        self.max_evals = int(maximum_number_of_function_evaluations)
        self.lower_bound = lower_bound
        self.upper_bound = upper_bound
        self.evaluations = 0

    def StartOptimization(self, do_only_initialization: bool = False):
        from scipy.optimize import minimize

        self._prepare_scales()
        if do_only_initialization:
            return
        metric = self.metric
        x0 = np.asarray(metric.GetParameters(), dtype=np.float64)
        self.evaluations = 0
        self.current_iteration = 0

        def fun(x):
            metric.SetParameters(x)
            v, d = metric.GetValueAndDerivative()
            self.evaluations += 1
            self.current_value = v
            return v, -np.asarray(d, dtype=np.float64)

        def callback(xk):
            self.current_iteration += 1
            for fn in self.observers:
                fn(self)

        bounds = None
        if self.lower_bound is not None or self.upper_bound is not None:
            lo = self.lower_bound if self.lower_bound is not None else -np.inf
            hi = self.upper_bound if self.upper_bound is not None else np.inf
            bounds = list(zip(np.broadcast_to(lo, x0.shape), np.broadcast_to(hi, x0.shape)))
        eps = np.finfo(np.float64).eps
        res = minimize(fun, x0, jac=True, method="L-BFGS-B", bounds=bounds, callback=callback,
                       options={"maxiter": self.number_of_iterations, "maxfun": self.max_evals,
                                "maxcor": self.m, "ftol": self.factr * eps, "gtol": self.pgtol})
        metric.SetParameters(res.x)
        self.current_value = float(res.fun)
        self.stop_description = "LBFGSBOptimizerv4: " + str(res.message)
        self.stop_condition = (StopCondition.MAXIMUM_NUMBER_OF_ITERATIONS if res.nit >= self.number_of_iterations
                               else StopCondition.CONVERGENCE_CHECKER_PASSED)
        self.result = res
