"""CoulombSummation and ChargeRelaxation: the pysic objects of the Ewald / EEM
path (pysic/interactions/coulomb.py, pysic/charges/relaxation.py)."""
import numpy as np


class InvalidCoulombError(Exception):
    pass


class InvalidRelaxationError(Exception):
    pass


class CoulombSummation:
    """Parameters of the Ewald summation (coulomb.py:6-176).

    parameters: real_cutoff (A), k_cutoff (1/A), sigma (A; the Gaussian
    width, erfc(r / (sqrt(2) sigma))), epsilon (e^2 / (eV A)).
    """

    summation_modes = ["ewald"]
    summation_parameters = {"ewald": ["real_cutoff", "k_cutoff", "sigma", "epsilon"]}

    def __init__(self, method="ewald", parameters=None, scaler=None):
        if method not in self.summation_modes:
            raise InvalidCoulombError("unknown summation mode: %s" % method)
        self.method = method
        self.parameters = {k: 0.0 for k in self.summation_parameters[method]}
        if parameters is not None:
            self.set_parameters(parameters)
        self.scaler = None if scaler is None else list(scaler)

    def set_parameters(self, parameters):
        if isinstance(parameters, dict):
            for k, v in parameters.items():
                self.set_parameter_value(k, v)
        else:
            names = self.summation_parameters[self.method]
            if len(parameters) != len(names):
                raise InvalidCoulombError("expected parameters %s" % names)
            for k, v in zip(names, parameters):
                self.parameters[k] = float(v)

    def set_parameter_value(self, name, value):
        if name not in self.parameters:
            raise InvalidCoulombError("no parameter %s" % name)
        self.parameters[name] = float(value)

    def get_parameters(self):
        return [self.parameters[k] for k in self.summation_parameters[self.method]]

    def get_summation(self):
        return self.method

    def set_scaling_factors(self, scaler):
        self.scaler = None if scaler is None else list(scaler)

    def get_scaling_factors(self):
        return self.scaler

    def get_realspace_cutoff(self):
        return self.parameters["real_cutoff"]

    def __eq__(self, other):
        return (isinstance(other, CoulombSummation) and self.method == other.method
                and self.parameters == other.parameters and self.scaler == other.scaler)


def estimate_ewald_parameters(real_cutoff=10.0, accuracy="normal"):
    """(real_cutoff, k_cutoff, sigma, epsilon) by the scaling rules of
    coulomb.py:179-218 (sigma ~ r_cut, k_cut ~ 1/sigma)."""
    epsilon = 0.00552635
    table = {"low": (2.0, 2.0), "normal": (3.0, 3.0), "high": (4.0, 4.0), "real": (3.0, 2.0),
             "reciprocal": (2.0, 3.0)}
    if accuracy not in table:
        raise InvalidCoulombError("accuracy must be one of %s" % sorted(table))
    ds, kk = table[accuracy]
    sigma = real_cutoff / ds
    return (real_cutoff, kk / sigma, sigma, epsilon)


class ChargeRelaxation:
    """Charge relaxation (relaxation.py:10-551).  The 'dynamic' mode (damped
    charge dynamics, relaxation.py:417-450) runs entirely on the device; the
    rates are carried across calls as in the reference (relaxation.py:431).

    An optional initial guess (a callable atoms -> charges, or a
    pysic_amd.learned.ChargePredictor, which predicts on the calculator's
    device-resident system) warm-starts the relaxation; the unchanged
    dynamics then runs to the same tolerance.
    """

    relaxation_modes = ["dynamic", "potentiostat", "optimize"]
    relaxation_parameters = {
        "dynamic": ["n_steps", "timestep", "inertia", "friction", "tolerance"],
        "potentiostat": ["n_steps", "timestep", "inertia", "potential"],
        "optimize": ["n_steps", "tolerance"],
    }

    def __init__(self, relaxation="dynamic", calculator=None, parameters=None, atoms=None):
        self.set_relaxation(relaxation)
        self.calculator = calculator
        self.parameters = {k: 0.0 for k in self.relaxation_parameters[self.relaxation]}
        if parameters is not None:
            self.set_parameters(parameters)
        self.atoms = atoms
        self.charge_rates = None
        self.observers = []
        self.initial_guess = None
        self.last_steps = 0
        self.last_error = None

    def set_relaxation(self, relaxation):
        if relaxation not in self.relaxation_modes:
            raise InvalidRelaxationError("unknown relaxation mode: %s" % relaxation)
        self.relaxation = relaxation

    def set_calculator(self, calculator, reciprocal=False):
        self.calculator = calculator
        if reciprocal and calculator is not None:
            calculator.set_charge_relaxation(self)

    def get_calculator(self):
        return self.calculator

    def set_atoms(self, atoms, pass_to_calculator=False):
        self.atoms = atoms
        if pass_to_calculator and self.calculator is not None:
            self.calculator.set_atoms(atoms)

    def set_parameters(self, parameters):
        if isinstance(parameters, dict):
            for k, v in parameters.items():
                self.set_parameter_value(k, v)
        else:
            names = self.relaxation_parameters[self.relaxation]
            for k, v in zip(names, parameters):
                self.parameters[k] = v

    def set_parameter_value(self, name, value):
        if name not in self.parameters:
            raise InvalidRelaxationError("no parameter %s" % name)
        self.parameters[name] = value

    def get_parameters(self):
        return dict(self.parameters)

    def set_initial_guess(self, predictor):
        """A callable atoms -> charges used to start the relaxation."""
        self.initial_guess = predictor

    def add_observer(self, function, interval=1, *args, **kwargs):
        """Observers are called once after the device loop ends with the
        number of steps taken (the loop runs without host round trips)."""
        self.observers.append((function, interval, args, kwargs))

    def call_observers(self, step):
        for fn, interval, args, kwargs in self.observers:
            if interval > 0 and step % interval == 0:
                fn(*args, **kwargs)

    def charge_relaxation(self):
        """Relax the charges of the calculator's structure; returns them."""
        if self.calculator is None:
            raise InvalidRelaxationError("no calculator attached")
        if self.relaxation != "dynamic":
            raise NotImplementedError("only the 'dynamic' relaxation runs on this engine")
        calc = self.calculator
        atoms = calc.get_atoms()
        charges = np.array(atoms.get_initial_charges(), dtype=np.float64)
        if self.initial_guess is not None:
            g = self.initial_guess
            if hasattr(g, "predict_for_calculator"):
                charges = np.array(g.predict_for_calculator(calc), dtype=np.float64)
            else:
                charges = np.array(g(atoms), dtype=np.float64)
        if self.charge_rates is None or len(self.charge_rates) != len(charges):
            self.charge_rates = np.zeros(len(charges))
        p = self.parameters
        q, rates, steps, err = calc.relax_charges(charges, self.charge_rates, int(p["n_steps"]),
                                                  float(p["timestep"]), float(p["inertia"]),
                                                  float(p["friction"]), float(p["tolerance"]))
        self.charge_rates = rates
        self.last_steps = steps
        self.last_error = err
        atoms.set_initial_charges(q)
        if self.atoms is not None and self.atoms is not atoms:
            self.atoms.set_initial_charges(q)
        self.call_observers(steps)
        return q
