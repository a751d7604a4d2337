class BasicQubit:
    def __init__(self, engine, idx):
        self.engine = engine
        self.id = idx

    def __int__(self):
        return int(self.engine.main_engine.get_measurement_result(self))

    def __bool__(self):
        return bool(int(self))

    def __eq__(self, other):
        return isinstance(other, BasicQubit) and self.id == other.id and self.id != -1

    def __hash__(self):
        return id(self)


class Qubit(BasicQubit):
    pass


class WeakQubitRef(BasicQubit):
    pass


class Qureg(list):
    pass
