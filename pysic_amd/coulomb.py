"""CoulombSummation and ChargeRelaxation: the pysic objects of the Ewald / EEM
path (pysic/interactions/coulomb.py, pysic/charges/relaxation.py)."""
import warnings

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
    summation_parameter_descriptions = {"ewald": [
        "cutoff of the real-space sum", "cutoff of the reciprocal-space sum",
        "width of the screening Gaussians (the real/reciprocal split)",
        "vacuum permittivity"]}

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
    relaxation_parameter_descriptions = {
        "dynamic": ["largest number of charge-dynamics steps", "time step of the charge dynamics",
                    "fictitious mass of the charges", "friction of the damped dynamics",
                    "stop when max(chi - mean chi) falls to this"],
        "potentiostat": ["number of charge-dynamics steps", "time step of the charge dynamics",
                         "fictitious mass of the charges", "potential of the external reservoir"],
        "optimize": ["largest number of optimizer iterations", "optimizer tolerance"],
    }

    def __init__(self, relaxation="dynamic", calculator=None, parameters=None, atoms=None):
        self.set_relaxation(relaxation)
        self.calculator = calculator
        self.parameters = {k: 0.0 for k in self.relaxation_parameters[self.relaxation]}
        if parameters is not None:
            self.set_parameters(parameters)
        self.atoms = atoms
        self.piped_relaxation = None
        self.charge_rates = None
        self.observers = []
        self.initial_guess = None
        self.last_steps = 0
        self.last_error = None

    def __eq__(self, other):
        """Same mode, calculator, parameters, atoms and pipe
        (relaxation.py:82-101)."""
        try:
            return (self.relaxation == other.relaxation and self.calculator == other.calculator
                    and self.parameters == other.parameters and self.atoms is other.atoms
                    and self.piped_relaxation == other.piped_relaxation)
        except Exception:
            return False

    def __ne__(self, other):
        return not self == other

    def __repr__(self):
        return "ChargeRelaxation('{m}',calculator={c},parameters={p},atoms={a})".format(
            m=self.relaxation, c=str(self.calculator), p=str(self.parameters), a=str(self.atoms))

    # -- pipes (relaxation.py:116-180, 300-322) ------------------------------
    def set_relaxation_pipe(self, relaxation, pass_info=True):
        """Run `relaxation` right after this one (and its own pipe after it).
        A pipe that would loop back to this relaxation is refused; with
        pass_info the calculator and atoms of this relaxation are handed down
        the pipe.  None clears the pipe."""
        if relaxation is None:
            self.piped_relaxation = None
            return
        if relaxation is self or any(r is self for r in relaxation.get_relaxation_pipe()):
            raise InvalidRelaxationError("Trying to create a looped relaxation pipe.")
        self.piped_relaxation = relaxation
        if pass_info:
            for rel in self.get_relaxation_pipe():
                rel.set_calculator(self.calculator)
                rel.set_atoms(self.atoms)
        if self.calculator != relaxation.calculator:
            warnings.warn("The piped ChargeRelaxation has a different calculator.", stacklevel=2)
        if self.atoms is not None and relaxation.atoms is None:
            warnings.warn("No Atoms object linked to the piped ChargeRelaxation.", stacklevel=2)

    def clear_relaxation_pipe(self):
        self.set_relaxation_pipe(None)

    def get_relaxation_pipe(self, recursive=True):
        """Every relaxation down the pipe (a list), or only the next one."""
        if not recursive:
            return self.piped_relaxation
        pipe = []
        r = self.piped_relaxation
        while r is not None:
            pipe.append(r)
            r = r.piped_relaxation
        return pipe

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

    def clear_observers(self):
        """Removes all observers (relaxation.py:533-537)."""
        self.observers = []

    def call_observers(self, step):
        for fn, interval, args, kwargs in self.observers:
            if interval > 0 and step % interval == 0:
                fn(*args, **kwargs)

    def charge_relaxation(self):
        """Relax the charges of the calculator's structure, then run the
        pipe; returns the final charges (relaxation.py:394-506).

        'dynamic' runs the damped dynamics on the device
        (relaxation.py:417-450); 'potentiostat' (relaxation.py:452-480) and
        'optimize' (relaxation.py:482-488, scipy's SLSQP with the total
        charge fixed) run the reference's host loops, with one device
        evaluation of the electronegativities (and energy) per step."""
        if self.calculator is None:
            raise InvalidRelaxationError("no calculator attached")
        calc = self.calculator
        atoms = calc.get_atoms()
        charges = np.array(atoms.get_initial_charges(), dtype=np.float64)
        if self.initial_guess is not None:
            g = self.initial_guess
            if hasattr(g, "predict_for_calculator"):
                charges = np.array(g.predict_for_calculator(calc), dtype=np.float64)
            else:
                charges = np.array(g(atoms), dtype=np.float64)
        p = self.parameters
        if self.relaxation == "dynamic":
            if self.charge_rates is None or len(self.charge_rates) != len(charges):
                self.charge_rates = np.zeros(len(charges))
            q, rates, steps, err = calc.relax_charges(
                charges, self.charge_rates, int(p["n_steps"]), float(p["timestep"]),
                float(p["inertia"]), float(p["friction"]), float(p["tolerance"]))
            self.charge_rates = rates
            self.last_steps = steps
            self.last_error = err
            atoms.set_initial_charges(q)
            self.call_observers(steps)
        elif self.relaxation == "potentiostat":
            q = self._potentiostat(calc, atoms, charges)
        else:
            q = self._optimize(calc, atoms, charges)
        if self.atoms is not None and self.atoms is not atoms:
            self.atoms.set_initial_charges(q)
        if self.piped_relaxation is not None:
            q = self.piped_relaxation.charge_relaxation()
        return q

    def _potentiostat(self, calc, atoms, charges):
        """Undamped charge dynamics against a reservoir at potential V:
        force chi - V, no mean removed (the total charge follows the
        reservoir), n_steps steps (relaxation.py:452-480)."""
        p = self.parameters
        dt = float(p["timestep"])
        inv_mq = 1.0 / float(p["inertia"])
        V = float(p["potential"])
        if self.charge_rates is None or len(self.charge_rates) != len(charges):
            self.charge_rates = np.zeros(len(charges))
        q = charges.copy()
        atoms.set_initial_charges(q)
        f = calc.get_electronegativities() - V
        for step in range(1, int(p["n_steps"]) + 1):
            q = q + self.charge_rates * dt + f * inv_mq * dt * dt
            self.charge_rates = self.charge_rates + f * 0.5 * inv_mq * dt
            atoms.set_initial_charges(q)
            f = calc.get_electronegativities() - V
            self.charge_rates = self.charge_rates + f * 0.5 * inv_mq * dt
            self.call_observers(step)
        self.last_steps = int(p["n_steps"])
        self.last_error = None
        return q

    def _optimize(self, calc, atoms, charges):
        """Minimise the energy over the charges at fixed total charge with
        scipy's SLSQP (relaxation.py:482-488, 373-392)."""
        from scipy.optimize import fmin_slsqp
        p = self.parameters
        total = float(np.sum(charges))

        def energy(q):
            atoms.set_initial_charges(q)
            return calc.get_potential_energy(skip_charge_relaxation=True)

        def gradient(q):
            atoms.set_initial_charges(q)
            return -calc.get_electronegativities()

        q = fmin_slsqp(energy, charges, fprime=gradient,
                       eqcons=[lambda q: float(np.sum(q)) - total],
                       acc=float(p["tolerance"]), iter=int(p["n_steps"]), disp=0)
        q = np.asarray(q, dtype=np.float64)
        atoms.set_initial_charges(q)
        self.last_steps = None
        self.last_error = None
        return q
