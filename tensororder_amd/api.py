"""The tensor-library plugin base (``tensor_network/tensor_apis/base_api.py:1-28``)."""

import itertools


class OutOfMemoryError(Exception):
    """Raised by a tensor library when the plan needs more slicing
    (``base_api.py:4-5``); ``execution.run`` then slices once more and retries."""


class BaseTensorAPI:
    def add_argument(self, key, value):
        raise ValueError("Invalid argument " + str(key) + " for selected tensor_library")

    def contract(self, network, contraction_tree):
        raise NotImplementedError

    def contract_sliced(self, execution_plan, num_slice_limit=None):
        """Sum over slices, in ``itertools.product`` order, of each slice's contraction."""
        result = 0
        slices = execution_plan.network.slice_groups(execution_plan.groups_to_slice)
        if num_slice_limit is not None:
            slices = itertools.islice(slices, num_slice_limit)
        for slice_network in slices:
            result += self.contract(slice_network, execution_plan.tree)[tuple()]
        return result
